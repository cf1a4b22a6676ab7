// mi_device.h -- device-side building blocks of the MI355X matrix-element integration path:
// Philox4x32-10 in registers, uniforms, this project's FP32 exp/log, the partonic gg->gg
// weight (FP64) and the hadronic 3-D integrand of the QCD 2->2 library in FP32 and FP64.
//
// Physics (paths relative to the reference repository):
//   sigmaHat brackets      pythia8317/src/SigmaQCD.cc:115-378 (Sigma2gg2gg ... Sigma2qqbar2qqbarNew)
//   CONVERT2MB             pythia8317/src/SigmaProcess.cc:24
//   parton flux + PDFs     pythia8317/src/SigmaProcess.cc:124-193 (initFlux), :423-480 (sigmaPDF)
//   (tau, y, z) map        pythia8317/src/PhaseSpace.cc:1256-1337 (selectTau, iTau = 0),
//                          :1373-1444 (y range), :1453-1568 (selectZ, stable tHat/uHat)
//   scales, alphaS         pythia8317/src/SigmaProcess.cc:876-922, pythia8317/src/StandardModel.cc:188-215
//   fused flat gg->gg      cuPythia/03_fused_resident_me/fused_gg2gg.cu:40-51
//
// The hadronic weight is written in terms of a = -t/s = (1 - z)/2 and b = -u/s = (1 + z)/2, both
// formed without cancellation from the uniforms, so FP32 holds ~1e-6 relative accuracy up to
// the pT cut.  With a, b, and L_x = sum over the flux list of x1 f(x1) x2 f(x2):
//   w_p = CONV pi LT / (4 pTmin^2) * alphaS^2 * (2 yMax) * zMax * (tauMin/tau) * L_p * B_p(a, b)
// which equals Pythia's sigmaPDF * wtTau * wtY * wtZ for tau = tauMin (1/tauMin)^u0,
// y = yMax (2 u1 - 1), z = zMax (2 u2 - 1).
//
// Heavy-quark pairs (codes 121..124, Sigma2gg2QQbar / Sigma2qqbar2QQbar, SigmaQCD.cc, with
// m3 = m4 = m): Pythia's massive phase space (tauMin = (2 mT)^2 / S with mT^2 = m^2 + pTmin^2,
// p2Abs = sH beta^2 / 4, z in [-zMax, zMax] with zMax = sqrt(1 - pTmin^2 / p2Abs)) is the
// massless map with pTmin^2 -> mT^2: beta z is uniform in [-sqrt(1 - 4 mT^2/sH), +...], so
// a = (m^2 - tH)/sH, b = (m^2 - uH)/sH and the Jacobian sH zMax' come out of the same formulas;
// Q2Ren = sqrt((pT^2 + m^2)^2) = Q2Fac = pT^2 + m^2 = sH a b.  Each family (massless, c cbar,
// b bbar) has its own map at the same uniforms, as each of Pythia's processes has its own
// PhaseSpace.
//
// FP32 arithmetic here is bit-reproducible on the host: only correctly rounded IEEE operations,
// explicit fmaf, no contraction (the library is compiled with -ffp-contract=off), and the own
// exp/log below (Cody-Waite reduction + fitted polynomial; table-driven log).  That is what makes
// the unweighting decisions, and hence the compacted event indices, bit-exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mi {

constexpr double kPi = 3.14159265358979323846;
constexpr double kConvert2mb = 0.389380;   // Pythia's GeV^-2 -> mb
constexpr double kLn2 = 0.69314718055994530942;
constexpr int kMaxProc = 10;      // slots 0..5 = codes 111..116, 6..9 = codes 121..124
constexpr int kLightProc = 6;
constexpr int kFamMask[3] = {0x3F, 0xC0, 0x300};   // proc_mask bits of each family

// ------------------------------------------------------------------ Philox4x32-10
struct U4 { uint32_t x, y, z, w; };

__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                             uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32 -> 64 multiply each (v_mad_u64_u32) instead of separate lo / hi products
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        // three-input XOR in one v_bitop3_b32 (truth table 0x96 = odd parity of the inputs)
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32(hi1, c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32(hi0, c3, k1, 0x96);
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    return U4{c0, c1, c2, c3};
}

// Event e of stream s under a 64-bit key: counter (e_lo, e_hi, s, 0), key (key_lo, key_hi).
__device__ __forceinline__ U4 philox_event(uint64_t key, uint64_t e, uint32_t stream)
{
    return philox4x32_10((uint32_t)e, (uint32_t)(e >> 32), stream, 0u, (uint32_t)key, (uint32_t)(key >> 32));
}

__device__ __forceinline__ double u53(uint32_t w0, uint32_t w1)
{
    const uint64_t x = ((uint64_t)w1 << 32) | (uint64_t)w0;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// 24-bit uniform as its integer k (u = k * 2^-24): exact in FP32 and FP64, and 1 - u is exact.
__device__ __forceinline__ uint32_t k24(uint32_t w) { return w >> 8; }

// SplitMix64 + u01 (reference-compatibility stream, cuPythia/common/rng.cuh:47-57).
__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double u01(uint64_t x) { return (double)(x >> 11) * (1.0 / 9007199254740992.0); }

// ------------------------------------------------------------------ own FP32 exp / log
// Branch-free forms (selects instead of branches, so a wave never runs both sides of a divergent
// test).  Every operation is correctly rounded (fma, mul, add, rint, ldexp, int <-> float), so
// the host restatement reproduces each result bit for bit.
//
// exp: Cody-Waite reduction x = n ln2 + r with ln2 split hi/lo, |r| <= ln2/2; e^r by a degree-6
// polynomial fitted for minimum relative error by Lawson's iteratively reweighted least squares
// (coefficients rounded to FP32 one by one, refitting the rest; relative error 3.1e-9);
// scaling by 2^n with ldexp.  Results that would leave the normal range are flushed to 0.
__device__ __forceinline__ float own_expf(float x)
{
    const bool under = x < -86.0f;
    const float xc = fminf(fmaxf(x, -86.0f), 88.0f);
    const float n = rintf(xc * 1.4426950408889634f);
    float r = __builtin_fmaf(n, -0.693145751953125f, xc);
    r = __builtin_fmaf(n, -1.4286068203094172e-6f, r);
    float p = 0.001382148f;
    p = __builtin_fmaf(p, r, 0.008368663f);
    p = __builtin_fmaf(p, r, 0.041668255f);
    p = __builtin_fmaf(p, r, 0.16666521f);
    p = __builtin_fmaf(p, r, 0.49999994f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    const float v = __builtin_ldexpf(p, (int)n);
    return under ? 0.0f : v;
}

// 1 - e^x (x <= 0) without cancellation near 0; ex = e^x already computed.  For x > -0.35:
// -x q(x) with q(x) ~ (e^x - 1)/x of degree 5 (same fitting method, relative error 2.5e-10);
// below that 1 - ex, which is >= 0.295 so the subtraction is benign.
__device__ __forceinline__ float one_minus_expf(float x, float ex)
{
    float p = __builtin_fmaf(x, 0.0011900244f, 0.008254852f);
    p = __builtin_fmaf(x, p, 0.041652385f);
    p = __builtin_fmaf(x, p, 0.16666552f);
    p = __builtin_fmaf(x, p, 0.49999997f);
    p = __builtin_fmaf(x, p, 1.0f);
    const float series = -(x * p);
    return (x > -0.35f) ? series : 1.0f - ex;
}

// log: table-driven reduction after P. T. P. Tang, "Table-driven implementation of the logarithm
// function in IEEE floating-point arithmetic", ACM TOMS 16 (1990) 378-400.  x = 2^e m, m in [1, 2);
// entry j = top 7 mantissa bits holds G_j = RN(1 / (1 + (j + 1/2)/128)) and
// L_j = RN(-log(G_j) - k_j ln2), k_j = 1 for j >= 53 (those entries fold m into [0.707, 1), so the
// result keeps relative accuracy just below x = 1); k_j is added to e by the carry of
// ix + (75 << 16).  r = m G_j - 1 (one fma, |r| <= 2^-8), log1p(r) ~ r - r^2/2 + r^3/3
// (truncation < 6e-11), log x = (e + k_j) ln2 + L_j + log1p(r).  No division.
static __constant__ float2 kLogTab[128] = {
    {0x1.fe01fep-1f, 0x1.ff00acp-9f}, {0x1.fa11cap-1f, 0x1.7dc49ep-7f},
    {0x1.f6310ap-1f, 0x1.3cea5ep-6f}, {0x1.f25f64p-1f, 0x1.b9fc0ap-6f},
    {0x1.ee9c8p-1f, 0x1.1b0d9p-5f}, {0x1.eae808p-1f, 0x1.58a5b6p-5f},
    {0x1.e741aap-1f, 0x1.95c836p-5f}, {0x1.e3a918p-1f, 0x1.d276b2p-5f},
    {0x1.e01e02p-1f, 0x1.075982p-4f}, {0x1.dca01ep-1f, 0x1.253f62p-4f},
    {0x1.d92f22p-1f, 0x1.42edcep-4f}, {0x1.d5cac8p-1f, 0x1.60658ap-4f},
    {0x1.d272cap-1f, 0x1.7da76ap-4f}, {0x1.cf26e6p-1f, 0x1.9ab422p-4f},
    {0x1.cbe6dap-1f, 0x1.b78c7ep-4f}, {0x1.c8b266p-1f, 0x1.d4313ap-4f},
    {0x1.c5894ep-1f, 0x1.f0a304p-4f}, {0x1.c26b54p-1f, 0x1.06715p-3f},
    {0x1.bf583ep-1f, 0x1.14785cp-3f}, {0x1.bc4fd6p-1f, 0x1.2266f4p-3f},
    {0x1.b951e2p-1f, 0x1.303d74p-3f}, {0x1.b65e2ep-1f, 0x1.3dfc2cp-3f},
    {0x1.b37484p-1f, 0x1.4ba372p-3f}, {0x1.b094b4p-1f, 0x1.59338ap-3f},
    {0x1.adbe88p-1f, 0x1.66acd4p-3f}, {0x1.aaf1d2p-1f, 0x1.740f94p-3f},
    {0x1.a82e66p-1f, 0x1.815c06p-3f}, {0x1.a5741p-1f, 0x1.8e929p-3f},
    {0x1.a2c2a8p-1f, 0x1.9bb366p-3f}, {0x1.a01a02p-1f, 0x1.a8becep-3f},
    {0x1.9d79f2p-1f, 0x1.b5b518p-3f}, {0x1.9ae24ep-1f, 0x1.c29688p-3f},
    {0x1.9852fp-1f, 0x1.cf635ap-3f}, {0x1.95cbbp-1f, 0x1.dc1bcep-3f},
    {0x1.934c68p-1f, 0x1.e8c026p-3f}, {0x1.90d4f2p-1f, 0x1.f550ap-3f},
    {0x1.8e6528p-1f, 0x1.00e6c4p-2f}, {0x1.8bfce8p-1f, 0x1.071b86p-2f},
    {0x1.899c1p-1f, 0x1.0d46b4p-2f}, {0x1.87427cp-1f, 0x1.13687p-2f},
    {0x1.84f00cp-1f, 0x1.1980d4p-2f}, {0x1.82a4ap-1f, 0x1.1f8ffap-2f},
    {0x1.806018p-1f, 0x1.259602p-2f}, {0x1.7e2256p-1f, 0x1.2b9302p-2f},
    {0x1.7beb3ap-1f, 0x1.31871ap-2f}, {0x1.79baa6p-1f, 0x1.377268p-2f},
    {0x1.779082p-1f, 0x1.3d54f8p-2f}, {0x1.756cacp-1f, 0x1.432ef2p-2f},
    {0x1.734f0cp-1f, 0x1.490068p-2f}, {0x1.713786p-1f, 0x1.4ec976p-2f},
    {0x1.6f2602p-1f, 0x1.548a2ap-2f}, {0x1.6d1a62p-1f, 0x1.5a42acp-2f},
    {0x1.6b149p-1f, 0x1.5ff308p-2f}, {0x1.691474p-1f, -0x1.602d0ap-2f},
    {0x1.6719f4p-1f, -0x1.5a8cbp-2f}, {0x1.6524f8p-1f, -0x1.54f43p-2f},
    {0x1.63356cp-1f, -0x1.4f638p-2f}, {0x1.614b36p-1f, -0x1.49da7ep-2f},
    {0x1.5f6644p-1f, -0x1.44592p-2f}, {0x1.5d867cp-1f, -0x1.3edf46p-2f},
    {0x1.5babccp-1f, -0x1.396ce2p-2f}, {0x1.59d62p-1f, -0x1.3401e4p-2f},
    {0x1.58056p-1f, -0x1.2e9e2cp-2f}, {0x1.56397cp-1f, -0x1.2941bp-2f},
    {0x1.54725ep-1f, -0x1.23ec58p-2f}, {0x1.52aff6p-1f, -0x1.1e9e18p-2f},
    {0x1.50f22ep-1f, -0x1.1956d4p-2f}, {0x1.4f38f6p-1f, -0x1.14167ep-2f},
    {0x1.4d843cp-1f, -0x1.0edd06p-2f}, {0x1.4bd3eep-1f, -0x1.09aa58p-2f},
    {0x1.4a27fap-1f, -0x1.047e5ep-2f}, {0x1.488052p-1f, -0x1.feb222p-3f},
    {0x1.46dce4p-1f, -0x1.f474b6p-3f}, {0x1.453d9ep-1f, -0x1.ea4448p-3f},
    {0x1.43a274p-1f, -0x1.e020d2p-3f}, {0x1.420b52p-1f, -0x1.d60a16p-3f},
    {0x1.40782ep-1f, -0x1.cc0012p-3f}, {0x1.3ee8f4p-1f, -0x1.c2028ap-3f},
    {0x1.3d5d9ap-1f, -0x1.b81178p-3f}, {0x1.3bd60ep-1f, -0x1.ae2caap-3f},
    {0x1.3a5244p-1f, -0x1.a4540cp-3f}, {0x1.38d22ep-1f, -0x1.9a877ep-3f},
    {0x1.3755bep-1f, -0x1.90c6e2p-3f}, {0x1.35dce6p-1f, -0x1.871214p-3f},
    {0x1.34679ap-1f, -0x1.7d68fep-3f}, {0x1.32f5cep-1f, -0x1.73cb8ap-3f},
    {0x1.318776p-1f, -0x1.6a39ap-3f}, {0x1.301c82p-1f, -0x1.60b30cp-3f},
    {0x1.2eb4eap-1f, -0x1.5737ccp-3f}, {0x1.2d50ap-1f, -0x1.4dc7b8p-3f},
    {0x1.2bef98p-1f, -0x1.4462b4p-3f}, {0x1.2a91cap-1f, -0x1.3b08bcp-3f},
    {0x1.293726p-1f, -0x1.31b996p-3f}, {0x1.27dfa4p-1f, -0x1.28753ep-3f},
    {0x1.268b38p-1f, -0x1.1f3b94p-3f}, {0x1.2539d8p-1f, -0x1.160c8p-3f},
    {0x1.23eb7ap-1f, -0x1.0ce7fp-3f}, {0x1.22a012p-1f, -0x1.03cdcp-3f},
    {0x1.215798p-1f, -0x1.f57bc8p-4f}, {0x1.201202p-1f, -0x1.e3708cp-4f},
    {0x1.1ecf44p-1f, -0x1.d1797cp-4f}, {0x1.1d8f56p-1f, -0x1.bf968p-4f},
    {0x1.1c523p-1f, -0x1.adc782p-4f}, {0x1.1b17c6p-1f, -0x1.9c0c2cp-4f},
    {0x1.19e012p-1f, -0x1.8a647ep-4f}, {0x1.18ab08p-1f, -0x1.78d02p-4f},
    {0x1.1778a2p-1f, -0x1.674f0ep-4f}, {0x1.1648d6p-1f, -0x1.55e10ep-4f},
    {0x1.151b9ap-1f, -0x1.4485dcp-4f}, {0x1.13f0e8p-1f, -0x1.333d74p-4f},
    {0x1.12c8b8p-1f, -0x1.2207acp-4f}, {0x1.11a302p-1f, -0x1.10e462p-4f},
    {0x1.107fbcp-1f, -0x1.ffa694p-5f}, {0x1.0f5eep-1f, -0x1.dda8b8p-5f},
    {0x1.0e4066p-1f, -0x1.bbced4p-5f}, {0x1.0d2446p-1f, -0x1.9a188ep-5f},
    {0x1.0c0a78p-1f, -0x1.78858ap-5f}, {0x1.0af2f8p-1f, -0x1.5715ep-5f},
    {0x1.09ddbap-1f, -0x1.35c8b2p-5f}, {0x1.08cabcp-1f, -0x1.149e56p-5f},
    {0x1.07b9f2p-1f, -0x1.e72bccp-6f}, {0x1.06ab5ap-1f, -0x1.a55f62p-6f},
    {0x1.059eeap-1f, -0x1.63d616p-6f}, {0x1.04949cp-1f, -0x1.228f82p-6f},
    {0x1.038c6cp-1f, -0x1.c3177cp-7f}, {0x1.02865p-1f, -0x1.4192bcp-7f},
    {0x1.018244p-1f, -0x1.8121bcp-8f}, {0x1.00804p-1f, -0x1.003fd6p-9f},
};

__device__ __forceinline__ float own_logf(float x)
{
    const bool bad = !(x > 0.0f);                          // only 1 - x == 0 reaches here
    const uint32_t ix = __float_as_uint(bad ? 1.0f : x);
    const float fe = (float)((ix + 0x004B0000u) >> 23) - 127.0f;
    const uint32_t j = (ix >> 16) & 127u;
    const float m = __uint_as_float((ix & 0x007fffffu) | 0x3f800000u);
    const float2 t = kLogTab[j];
    const float r = __builtin_fmaf(m, t.x, -1.0f);
    const float r2 = r * r;
    const float p = __builtin_fmaf(r2, __builtin_fmaf(r, 0.333333343f, -0.5f), r);
    const float u = t.y + __builtin_fmaf(fe, 1.4286068203094172e-6f, p);
    const float v = __builtin_fmaf(fe, 0.693145751953125f, u);
    return bad ? -1.0e30f : v;
}

// ------------------------------------------------------------------ FP32 math of the integrand
// HW = false: IEEE division and square root (with the own exp / log above: bit-reproducible on
// the host).  HW = true (MI_PREC_FP32_HW): the hardware's reciprocal and square root (v_rcp_f32,
// v_sqrt_f32, about 1 ulp each); that mode also takes the hardware's base-2 exponential and
// logarithm (v_exp_f32, v_log_f32) directly, in the base-2 form of the integrand.
template <bool HW>
__device__ __forceinline__ float frcp(float x)
{
    if constexpr (!HW) return 1.0f / x;
    else return __builtin_amdgcn_rcpf(x);
}
template <bool HW>
__device__ __forceinline__ float fsqrt(float x)
{
    if constexpr (!HW) return sqrtf(x);
    else return __builtin_amdgcn_sqrtf(x);
}

// 1 - 2^t (t <= 0) without cancellation near 0, for the base-2 form of MI_PREC_FP32_HW; e = 2^t
// already computed.  For t > -0.50494 (x = t ln 2 > -0.35): one_minus_expf's series
// -x q(x) written in t, i.e. its coefficients c_k times ln 2^(k+1); below, 1 - e >= 0.295.
__device__ __forceinline__ float one_minus_exp2f(float t, float e)
{
    constexpr double L = kLn2, L2 = L * L, L3 = L2 * L, L4 = L3 * L, L5 = L4 * L, L6 = L5 * L;
    float p = __builtin_fmaf(t, (float)(0.0011900244 * L6), (float)(0.008254852 * L5));
    p = __builtin_fmaf(t, p, (float)(0.041652385 * L4));
    p = __builtin_fmaf(t, p, (float)(0.16666552 * L3));
    p = __builtin_fmaf(t, p, (float)(0.49999997 * L2));
    p = __builtin_fmaf(t, p, (float)L);
    const float series = -(t * p);
    return (t > -0.50494f) ? series : 1.0f - e;
}

// ------------------------------------------------------------------ constants of the integrand
// Filled on the host by fill_consts (mi_api.hip).
template <typename T>
struct HadConst {
    T LT;          // ln(1 / tauMin)
    T fourPT2;     // 4 pTmin^2 = tauMin S
    T pref;        // CONVERT2MB pi LT K / (4 pTmin^2)
    T xiR2;        // renorm_fac^2
    T as_order;    // 0 / 1
    T as_fixed;
    T scale2Min, mt2, mb2, mc2, nf6;
    T lnL6s, lnL5s, lnL4s, lnL3s;
    T c6, c5, c4, c3;  // 12 pi / b0
    T lnAg, ag, bg, lnAs, as, bs, lnAuv, av, bv, rdu;
    T nSea;        // 2 nQuarkIn
    T nQuarkNew;
    T spare;
    // heavy-quark families 0 = c cbar (121, 122), 1 = b bbar (123, 124)
    struct Fam {
        T LT;      // ln(S / (4 mT^2)), mT^2 = m^2 + pTmin^2
        T fourMT2; // 4 mT^2
        T pref;    // CONVERT2MB pi LT K / (4 mT^2)
        T m2;      // m^2
        T mu;      // m^2 / (4 mT^2), so m^2 / sH = mu r
    } fam[2];
    // MI_PREC_FP32_HW evaluates the map, alphaS and the toy PDFs in base 2 (v_exp_f32 and
    // v_log_f32 take and return base-2 values): the constants above that sit in an exponent or
    // come out of a logarithm, divided by ln 2, and the Jacobian's prefactors times ln 2.
    struct Base2 {
        T LT;                      // LT / ln 2
        T pref;                    // pref ln 2: the Jacobian's LT om0 = ln 2 (LT / ln 2) om0
        T lnAg, lnAs, lnAuv;       // log2 of the PDF normalisations
        T lam6, lam5, lam4, lam3;  // ln(Lambda_nf^2) / ln 2
        T c6, c5, c4, c3;          // (12 pi / b0) / ln 2: alphaS = c / (log2 Q2 - lam)
        T famLT[2], famPref[2];    // the heavy families' LT / ln 2 and pref ln 2
    } b2;
    T nSeaM2, nSeaM4;              // nSea - 2, nSea - 4 (exact), so no kernel recomputes them per point
    T above_mb;                    // 1: renorm_fac^2 pTmin^2 > 1.01 mb^2, so every point's Q2 > mb^2
};
static_assert(sizeof(HadConst<float>) == 62 * sizeof(float), "HadConst layout");

// ------------------------------------------------------------------ PDF grid (LHAGrid1)
// A proton PDF read from an LHAPDF6 "lhagrid1" file (mi_pdf_load), evaluated as Pythia's
// LHAGrid1::xfxevolve does (pythia8317/src/PartonDistributions.cc:648-745, PDF:extrapolate =
// off): bisection for the neighbouring nodes in x and in q (within the q subgrid holding q; the
// FP32 device path starts from node-guess tables and steps to the same interval),
// 4-point Lagrange weights in ln x and ln q (linear in ln q on subgrids of 2-3 points), frozen
// outside the q range and below xMin, 0 at x >= xMax.  Slots: 0 = g, 1..5 = d u s c b,
// 6..10 = dbar ubar sbar cbar bbar, 11 = photon; values flavour-interleaved
// val[(iq * nx + ix) * 12 + slot].  FP64 weights divide as Pythia does; the FP32 weights
// multiply by reciprocals tabulated on the host (ixinv, iqinv: 12 per stencil start, the
// (i, j != i) pairs in order; iqinv[12 m] = 1 / (lnq[m+1] - lnq[m]) for 2-3 point subgrids).
constexpr int kGridNF = 12, kGridMaxSub = 8;

template <typename T>
struct GridPdf {
    const T *xk = nullptr, *lnx = nullptr;       // nx: nodes (search keys) and their logs
    const T *qk = nullptr, *lnq = nullptr;       // nq: nodes with subgrid borders averaged; logs as read
    const T *val = nullptr;                      // nq x nx x 12
    const T *ixinv = nullptr, *iqinv = nullptr;  // FP32 only
    // FP32 only: node guesses on a uniform ln x (ln q) binning, xlut[j] = the last node at or
    // below bin j's left edge; the search then steps with the bisection's own comparisons
    const unsigned short *xlut = nullptr, *qlut = nullptr;
    float xl0 = 0, xih = 0, ql0 = 0, qih = 0;
    int nbx = 0, nbq = 0;
    // FP32 only: a window of val (rows wq0..wq1, columns wx0..nx-1, wstride = nx - wx0 nodes per
    // row) copied to the kernel's dynamic LDS at byte offset woff (-1: no window)
    int woff = -1, wq0 = 0, wq1 = -1, wx0 = 0, wstride = 0;
    T qDiv[kGridMaxSub] = {};                    // last q of each subgrid, as read
    T xMin = 0, xMax = 0, qMin = 0, qMax = 0;
    T factFac2 = 1;                              // Q2_F = factFac2 * pT2
    int nqSum[kGridMaxSub] = {};                 // cumulative q count after each subgrid
    int nx = 0, nq = 0, nqSub = 0, active = 0;
};

// w[i] = prod_{j != i} (t - n[j]) / (n[i] - n[j]) over the 4 nodes n (Lagrange weights).
template <typename T>
__host__ __device__ inline void lagrange4(T t, const T *n, const T *inv, T *w)
{
    int k = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        w[i] = (T)1;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j != i) {
                if (inv) w[i] *= (t - n[j]) * inv[k];
                else     w[i] *= (t - n[j]) / (n[i] - n[j]);
                ++k;
            }
    }
}

// q stencil shared by both partons: start row m3q, n3q rows (1, 2 or 4) with weights wq.
template <typename T>
__host__ __device__ inline void grid_qw(const GridPdf<T> &g, T q, T lnq, int &m3q, int &n3q, T *wq)
{
    wq[0] = wq[1] = wq[2] = wq[3] = (T)1;
    if (q <= g.qMin) { n3q = 1; m3q = 0; return; }
    if (q >= g.qMax) { n3q = 1; m3q = g.nq - 1; return; }
    int iqDiv = 0;
    for (int s = 1; s < g.nqSub; ++s)
        if (q > g.qDiv[s - 1]) iqDiv = s;
    const int minS = iqDiv == 0 ? 0 : g.nqSum[iqDiv - 1], maxS = g.nqSum[iqDiv] - 1;
    int minq = minS, maxq = maxS;
    while (maxq - minq > 1) {
        const int mid = (minq + maxq) >> 1;
        if (q < g.qk[mid]) maxq = mid;
        else               minq = mid;
    }
    if (maxS - minS < 3) {
        n3q = 2;
        m3q = minq;
        wq[1] = g.iqinv ? (lnq - g.lnq[minq]) * g.iqinv[kGridNF * minq] : (lnq - g.lnq[minq]) / (g.lnq[maxq] - g.lnq[minq]);
        wq[0] = (T)1 - wq[1];
    } else {
        n3q = 4;
        m3q = (minq == minS) ? minS : ((maxq == maxS) ? maxS - 3 : minq - 1);
        lagrange4<T>(lnq, g.lnq + m3q, g.iqinv ? g.iqinv + kGridNF * m3q : nullptr, wq);
    }
}

// Raw interpolated xf of slots 0..10 at (x, ln x) for the q stencil of grid_qw.
template <typename T>
__host__ __device__ inline void grid_xf(const GridPdf<T> &g, T x, T lnx, int m3q, int n3q, const T *wq, T *out)
{
#pragma unroll
    for (int f = 0; f < 11; ++f) out[f] = (T)0;
    if (x <= g.xMin) {
        for (int i3q = 0; i3q < n3q; ++i3q) {
            const T *p = g.val + (size_t)(m3q + i3q) * g.nx * kGridNF;
#pragma unroll
            for (int f = 0; f < 11; ++f) out[f] += wq[i3q] * p[f];
        }
        return;
    }
    if (x >= g.xMax) return;
    int minx = 0, maxx = g.nx - 1;
    while (maxx - minx > 1) {
        const int mid = (minx + maxx) >> 1;
        if (x < g.xk[mid]) maxx = mid;
        else               minx = mid;
    }
    const int m3x = (minx == 0) ? 0 : ((maxx == g.nx - 1) ? g.nx - 4 : minx - 1);
    T wx[4];
    lagrange4<T>(lnx, g.lnx + m3x, g.ixinv ? g.ixinv + kGridNF * m3x : nullptr, wx);
    for (int i3q = 0; i3q < n3q; ++i3q) {
        const T *p = g.val + ((size_t)(m3q + i3q) * g.nx + m3x) * kGridNF;
#pragma unroll
        for (int f = 0; f < 11; ++f)
            out[f] += wq[i3q] * (wx[0] * p[f] + wx[1] * p[kGridNF + f] + wx[2] * p[2 * kGridNF + f]
                                 + wx[3] * p[3 * kGridNF + f]);
    }
}

// The dynamic LDS of the running kernel (the integration kernel keeps its PDF-grid window there).
extern __shared__ __align__(16) unsigned char mi_lds[];

// FP32 (device layout): the same arithmetic with each node's 12 values and each stencil start's
// 12 reciprocals fetched as three 16-byte words (the FP32 tables are 16-byte aligned).
__host__ __device__ inline void lagrange4_v(float t, const float *n, const float *inv, float *w)
{
    const float4 *iv = reinterpret_cast<const float4 *>(inv);
    const float4 a = iv[0], b = iv[1], c = iv[2];
    const float r[12] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w};
    int k = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        w[i] = 1.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j != i) { w[i] *= (t - n[j]) * r[k]; ++k; }
    }
}

__host__ __device__ inline void node_vals(const float *p, float *v)
{
    const float4 *q = reinterpret_cast<const float4 *>(p);
    const float4 a = q[0], b = q[1], c = q[2];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y;
    v[6] = b.z; v[7] = b.w; v[8] = c.x; v[9] = c.y; v[10] = c.z;
}

// Bin of a log value in a node-guess table (clamped to the table).
__host__ __device__ inline int lut_bin(float l, float l0, float ih, int nb)
{
    const float t = (l - l0) * ih;
    return t <= 0.0f ? 0 : (t >= (float)(nb - 1) ? nb - 1 : (int)t);
}

__device__ inline void grid_qw_f32(const GridPdf<float> &g, float q, float lnq, int &m3q, int &n3q, float *wq)
{
    wq[0] = wq[1] = wq[2] = wq[3] = 1.0f;
    if (q <= g.qMin) { n3q = 1; m3q = 0; return; }
    if (q >= g.qMax) { n3q = 1; m3q = g.nq - 1; return; }
    // the subgrid holding q (the last one whose predecessor ends below q), with constant
    // indices only, so the view's arrays stay in scalar registers
    int minS = 0, maxS = g.nqSum[0] - 1;
#pragma unroll
    for (int s = 1; s < kGridMaxSub; ++s)
        if (s < g.nqSub && q > g.qDiv[s - 1]) { minS = g.nqSum[s - 1]; maxS = g.nqSum[s] - 1; }
    // the interval the bisection over [minS, maxS] finds: the last node in [minS, maxS - 1] at or
    // below q (minS if none), from the table's guess in a few steps
    int minq = g.qlut[lut_bin(lnq, g.ql0, g.qih, g.nbq)];
    minq = minq < minS ? minS : (minq > maxS - 1 ? maxS - 1 : minq);
    while (minq > minS && q < g.qk[minq]) --minq;
    while (minq < maxS - 1 && q >= g.qk[minq + 1]) ++minq;
    const int maxq = minq + 1;
    if (maxS - minS < 3) {
        n3q = 2;
        m3q = minq;
        wq[1] = (lnq - g.lnq[minq]) * g.iqinv[kGridNF * minq];
        wq[0] = 1.0f - wq[1];
    } else {
        n3q = 4;
        m3q = (minq == minS) ? minS : ((maxq == maxS) ? maxS - 3 : minq - 1);
        lagrange4_v(lnq, g.lnq + m3q, g.iqinv + kGridNF * m3q, wq);
    }
}

// Folds n3q rows of 4 nodes (row stride rs floats from p) with the combined weights wq * wx.
template <typename P>
__device__ __forceinline__ void fold_stencil(P p, size_t rs, int n3q, const float *wq, const float *wx, float *out)
{
#pragma unroll
    for (int i3q = 0; i3q < 4; ++i3q) {             // n3q <= 4 rows; constant indices keep wq in registers
        if (i3q >= n3q) continue;
#pragma unroll
        for (int ix = 0; ix < 4; ++ix) {
            const float w = wq[i3q] * wx[ix];
            float v[11];
            node_vals(p + i3q * rs + ix * kGridNF, v);
#pragma unroll
            for (int f = 0; f < 11; ++f) out[f] = __builtin_fmaf(w, v[f], out[f]);
        }
    }
}

// The 4 x 4 stencil (n3q rows of 4 x nodes) folded node by node with the combined weight
// wq * wx into FP32 fmas: 16 products and 11 fmas per node, one node's values live at a time.
// Node values come from the LDS window when the stencil lies inside it, else from global memory
// (the same values, the same order of operations).
__device__ inline void grid_xf_f32(const GridPdf<float> &g, float x, float lnx, int m3q, int n3q,
                                   const float *wq, float *out)
{
#pragma unroll
    for (int f = 0; f < 11; ++f) out[f] = 0.0f;
    if (x <= g.xMin) {
#pragma unroll
        for (int i3q = 0; i3q < 4; ++i3q) {
            if (i3q >= n3q) continue;
            float v[11];
            const bool lds = g.woff >= 0 && g.wx0 == 0 && m3q + i3q >= g.wq0 && m3q + i3q <= g.wq1;
            if (lds)
                node_vals(reinterpret_cast<const float *>(mi_lds + g.woff) +
                              (size_t)(m3q + i3q - g.wq0) * g.wstride * kGridNF, v);
            else
                node_vals(g.val + (size_t)(m3q + i3q) * g.nx * kGridNF, v);
#pragma unroll
            for (int f = 0; f < 11; ++f) out[f] = __builtin_fmaf(wq[i3q], v[f], out[f]);
        }
        return;
    }
    if (x >= g.xMax) return;
    // xk[minx] <= x < xk[minx + 1] (xMin < x < xMax), as the bisection finds it
    int minx = g.xlut[lut_bin(lnx, g.xl0, g.xih, g.nbx)];
    while (minx > 0 && x < g.xk[minx]) --minx;
    while (minx < g.nx - 2 && x >= g.xk[minx + 1]) ++minx;
    const int maxx = minx + 1;
    const int m3x = (minx == 0) ? 0 : ((maxx == g.nx - 1) ? g.nx - 4 : minx - 1);
    float wx[4];
    lagrange4_v(lnx, g.lnx + m3x, g.ixinv + kGridNF * m3x, wx);
    if (g.woff >= 0 && m3x >= g.wx0 && m3q >= g.wq0 && m3q + n3q - 1 <= g.wq1)
        fold_stencil(reinterpret_cast<const float *>(mi_lds + g.woff) +
                         ((size_t)(m3q - g.wq0) * g.wstride + (m3x - g.wx0)) * kGridNF,
                     (size_t)g.wstride * kGridNF, n3q, wq, wx, out);
    else
        fold_stencil(g.val + ((size_t)m3q * g.nx + m3x) * kGridNF, (size_t)g.nx * kGridNF, n3q, wq, wx, out);
}

// Parton luminosities of the five flux types from the grid at the factorization scale
// (PDF values clipped at 0 as Pythia's PDF::xf does): gg, qg, all quark pairs QQ, same-flavour
// pairs (q q and qbar qbar), and q qbar pairs of one flavour; nQ flavours enter.
template <typename T>
struct Lumi { T gg, qg, QQ, same, opp, g1, g2, Q1, Q2; };

template <typename T>
__host__ __device__ inline Lumi<T> grid_lumi(const GridPdf<T> &g, int nQ, T x1, T lx1, T x2, T lx2, T q, T lnq)
{
    int m3q, n3q;
    T wq[4], f1[11], f2[11];
    if constexpr (sizeof(T) == 4) {
        grid_qw_f32(g, q, lnq, m3q, n3q, wq);
        grid_xf_f32(g, x1, lx1, m3q, n3q, wq, f1);
        grid_xf_f32(g, x2, lx2, m3q, n3q, wq, f2);
    } else {
        grid_qw<T>(g, q, lnq, m3q, n3q, wq);
        grid_xf<T>(g, x1, lx1, m3q, n3q, wq, f1);
        grid_xf<T>(g, x2, lx2, m3q, n3q, wq, f2);
    }
#pragma unroll
    for (int f = 0; f < 11; ++f) {
        f1[f] = f1[f] > (T)0 ? f1[f] : (T)0;
        f2[f] = f2[f] > (T)0 ? f2[f] : (T)0;
    }
    Lumi<T> L;
    T Q1 = 0, Q2 = 0, same = 0, opp = 0;
    // a fixed trip count with the flavours beyond nQ skipped: constant indices keep f1, f2 in
    // registers
#pragma unroll
    for (int k = 1; k <= 5; ++k) {
        if (k > nQ) continue;
        const T q1 = f1[k], a1 = f1[5 + k], q2 = f2[k], a2 = f2[5 + k];
        Q1 = Q1 + (q1 + a1);
        Q2 = Q2 + (q2 + a2);
        same = same + (q1 * q2 + a1 * a2);
        opp = opp + (q1 * a2 + a1 * q2);
    }
    L.g1 = f1[0]; L.g2 = f2[0]; L.Q1 = Q1; L.Q2 = Q2;
    L.gg = f1[0] * f2[0];
    L.qg = Q1 * f2[0] + f1[0] * Q2;
    L.QQ = Q1 * Q2;
    L.same = same;
    L.opp = opp;
    return L;
}

// Intermediate values the adjoint pass needs from a point (see mi_adjoint.hip).
template <typename T>
struct PointAux { T alpS, cf, lx1, lx2, l1a, l1b, g1, g2, Q1, Q2; };

// ------------------------------------------------------------------ FP32 hadronic point
// One family (FAM 0: the massless codes 111..116 into ws[0..5]; 1, 2: c cbar, b bbar into
// ws[6..7], ws[8..9]).  Inputs: (u, 1-u) for each dimension (exact pairs); writes the family's
// slots (0 if masked out) and returns their sum.  HW = false: operation-for-operation identical
// to the host restatement used to check compaction indices.  HW = true (MI_PREC_FP32_HW): the
// map, alphaS and the toy PDFs in base 2, so that v_exp_f32 / v_log_f32 take and return the
// values directly (log2 tau, log2 x, log2 (1 - x), log2 Q2 and the PDF exponents; HadConst::b2);
// the grid PDFs still receive ln x, computed as the bit-exact path does.  aux: HW = false only.
// U24 (HW only): u and 1 - u arrive as the integers k and 2^24 - k of the 24-bit uniforms, and
// the 2^-24 goes into the products that take them (exact scalings).  BR: keep the massless
// slots' mask tests as branches (off where the three families share a point: registers).
template <int FAM, bool HW = false, bool U24 = false, bool BR = true>
__device__ __forceinline__ float had_family_f32(const HadConst<float>& k, int mask,
                                                float u0, float om0, float u1, float om1,
                                                float u2, float om2, float* ws, float* sHo, float* tHo,
                                                float* x1o, float* x2o, PointAux<float>* aux,
                                                const GridPdf<float>* gp, bool grid)
{
    constexpr int F = FAM > 0 ? FAM - 1 : 0;                 // heavy family index
    const float LT = FAM == 0 ? k.LT : k.fam[F].LT;
    float r, jacY, lx1, lx2, x1, x2, zMax;                 // jacY: 2 yMax in the prefactor's units
    static_assert(HW || !U24, "integer uniforms are for the base-2 form");
    constexpr float su = U24 ? 0x1p-24f : 1.0f;            // the uniforms' unit
    if constexpr (HW) {
        const float LT2 = (FAM == 0 ? k.b2.LT : k.b2.famLT[F]) * su;
        const float targ = -(u0 * LT2);                    // log2(tauMin / tau)
        r = __builtin_amdgcn_exp2f(targ);                  // tauMin / tau
        jacY = LT2 * om0;                                  // 2 yMax / ln 2
        const float jy = jacY * su;
        lx1 = -(jy * om1); lx2 = -(jy * u1);               // log2 x1, log2 x2
        x1 = __builtin_amdgcn_exp2f(lx1); x2 = __builtin_amdgcn_exp2f(lx2);
        zMax = __builtin_amdgcn_sqrtf(one_minus_exp2f(targ, r));
    } else {
        const float targ = -(u0 * LT);
        r = own_expf(targ);                             // tauMin / tau
        const float yMax = 0.5f * (LT * om0);
        const float twoYMax = yMax + yMax;
        lx1 = -(twoYMax * om1); lx2 = -(twoYMax * u1);
        x1 = own_expf(lx1); x2 = own_expf(lx2);
        zMax = fsqrt<false>(one_minus_expf(targ, r));
        jacY = LT * om0;
    }
    // one division gives both 1 / r and 1 / (1 + zMax)
    const float opz = 1.0f + zMax;
    const float R1 = frcp<HW>(r * opz);
    const float omz = (r * r) * R1;                        // r / (1 + zMax) = 1 - zMax
    const float sH = ((FAM == 0 ? k.fourPT2 : k.fam[F].fourMT2) * opz) * R1;   // 4 mT^2 / r
    const float twoz = U24 ? zMax * (2.0f * su) : zMax + zMax;
    const float A = __builtin_fmaf(twoz, om2, omz);        // 1 - z
    const float B = __builtin_fmaf(twoz, u2, omz);         // 1 + z
    const float a = 0.5f * A, b = 0.5f * B;                // (m^2 - tH) / sH, (m^2 - uH) / sH
    const float ab = a * b;
    const float pT2 = sH * ab;                             // pT^2 + m^2

    float alpS, inv_ab, cfo = 0.0f;
    if (k.as_order == 0.0f) {
        alpS = k.as_fixed;
        inv_ab = frcp<HW>(ab);
    } else {
        float Q2 = k.xiR2 * pT2;
        Q2 = (Q2 < k.scale2Min) ? k.scale2Min : Q2;
        const bool t6 = Q2 > k.mt2 && k.nf6 != 0.0f, t5 = Q2 > k.mb2, t4 = Q2 > k.mc2;
        float lq, lam, cf;
        // HW: base 2 (Q2 >= scale2Min > 0).  above_mb (uniform): Q2 > mb^2 at every point of the
        // run, so the flavour number is 5 or 6 and the chain shortens to one test
        const float L6 = HW ? k.b2.lam6 : k.lnL6s, L5 = HW ? k.b2.lam5 : k.lnL5s;
        const float C6 = HW ? k.b2.c6 : k.c6, C5 = HW ? k.b2.c5 : k.c5;
        if (k.above_mb != 0.0f) {
            lam = t6 ? L6 : L5;
            cf = t6 ? C6 : C5;
        } else if constexpr (HW) {
            lam = t6 ? L6 : (t5 ? L5 : (t4 ? k.b2.lam4 : k.b2.lam3));
            cf = t6 ? C6 : (t5 ? C5 : (t4 ? k.b2.c4 : k.b2.c3));
        } else {
            lam = t6 ? L6 : (t5 ? L5 : (t4 ? k.lnL4s : k.lnL3s));
            cf = t6 ? C6 : (t5 ? C5 : (t4 ? k.c4 : k.c3));
        }
        if constexpr (HW) lq = __builtin_amdgcn_logf(Q2);
        else lq = own_logf(Q2);
        // one division gives both 1 / (a b) and alphaS = cf / (lq - lam)
        const float den = lq - lam;
        const float R2 = frcp<HW>(ab * den);
        inv_ab = den * R2;
        alpS = (cf * ab) * R2;
        cfo = cf;
    }
    const float ia = b * inv_ab, ib = a * inv_ab;

    float Lgg, Lqg, LQQ, Lsame, Lopp, l1a = 0.0f, l1b = 0.0f, g1, g2, Q1, Q2s;
    if (grid) {
        // PDF grid at Q2_F = factFac^2 pT2 (q = sqrt, ln q = ln Q2_F / 2)
        const float Q2F = gp->factFac2 * pT2;
        float lnx1 = lx1, lnx2 = lx2, lnq;
        if constexpr (HW) {
            const float twoYMax = (LT * su) * om0;         // ln x as the bit-exact path forms it
            lnx1 = -((twoYMax * su) * om1); lnx2 = -((twoYMax * su) * u1);
            lnq = (float)(0.5 * kLn2) * __builtin_amdgcn_logf(Q2F);
        } else {
            lnq = 0.5f * own_logf(Q2F);
        }
        const Lumi<float> L = grid_lumi<float>(*gp, (int)(0.5f * k.nSea), x1, lnx1, x2, lnx2, fsqrt<HW>(Q2F), lnq);
        Lgg = L.gg; Lqg = L.qg; LQQ = L.QQ; Lsame = L.same; Lopp = L.opp;
        g1 = L.g1; g2 = L.g2; Q1 = L.Q1; Q2s = L.Q2;
    } else {
        float S1, S2, V1, V2;
        if constexpr (HW) {                                // log2 xf = log2 A + a log2 x + b log2 (1 - x)
            l1a = __builtin_amdgcn_logf(one_minus_exp2f(lx1, x1));
            l1b = __builtin_amdgcn_logf(one_minus_exp2f(lx2, x2));
            g1 = __builtin_amdgcn_exp2f(__builtin_fmaf(k.bg, l1a, __builtin_fmaf(k.ag, lx1, k.b2.lnAg)));
            g2 = __builtin_amdgcn_exp2f(__builtin_fmaf(k.bg, l1b, __builtin_fmaf(k.ag, lx2, k.b2.lnAg)));
            S1 = __builtin_amdgcn_exp2f(__builtin_fmaf(k.bs, l1a, __builtin_fmaf(k.as, lx1, k.b2.lnAs)));
            S2 = __builtin_amdgcn_exp2f(__builtin_fmaf(k.bs, l1b, __builtin_fmaf(k.as, lx2, k.b2.lnAs)));
            V1 = __builtin_amdgcn_exp2f(__builtin_fmaf(k.bv, l1a, __builtin_fmaf(k.av, lx1, k.b2.lnAuv)));
            V2 = __builtin_amdgcn_exp2f(__builtin_fmaf(k.bv, l1b, __builtin_fmaf(k.av, lx2, k.b2.lnAuv)));
        } else {
            l1a = own_logf(one_minus_expf(lx1, x1)); l1b = own_logf(one_minus_expf(lx2, x2));
            g1 = own_expf(__builtin_fmaf(k.bg, l1a, __builtin_fmaf(k.ag, lx1, k.lnAg)));
            g2 = own_expf(__builtin_fmaf(k.bg, l1b, __builtin_fmaf(k.ag, lx2, k.lnAg)));
            S1 = own_expf(__builtin_fmaf(k.bs, l1a, __builtin_fmaf(k.as, lx1, k.lnAs)));
            S2 = own_expf(__builtin_fmaf(k.bs, l1b, __builtin_fmaf(k.as, lx2, k.lnAs)));
            V1 = own_expf(__builtin_fmaf(k.bv, l1a, __builtin_fmaf(k.av, lx1, k.lnAuv)));
            V2 = own_expf(__builtin_fmaf(k.bv, l1b, __builtin_fmaf(k.av, lx2, k.lnAuv)));
        }
        const float Vd1 = k.rdu * V1, Vd2 = k.rdu * V2;
        const float U1 = V1 + S1, U2 = V2 + S2, D1 = Vd1 + S1, D2 = Vd2 + S2;
        Q1 = __builtin_fmaf(k.nSea, S1, V1 + Vd1); Q2s = __builtin_fmaf(k.nSea, S2, V2 + Vd2);
        const float SS = S1 * S2;
        Lgg = g1 * g2;
        Lqg = __builtin_fmaf(Q1, g2, g1 * Q2s);
        LQQ = Q1 * Q2s;
        Lsame = __builtin_fmaf(U1, U2, __builtin_fmaf(D1, D2, k.nSeaM2 * SS));
        Lopp = __builtin_fmaf(U1 + D1, S2, __builtin_fmaf(S1, U2 + D2, k.nSeaM4 * SS));
    }

    const float q2 = __builtin_fmaf(a, a, b * b);
    const float pref = HW ? (FAM == 0 ? k.b2.pref : k.b2.famPref[F]) : (FAM == 0 ? k.pref : k.fam[F].pref);
    const float common = pref * (alpS * alpS) * jacY * zMax * r;
    float tot = 0.0f;
    if (FAM == 0) {
        float w[6];
        w[0] = Lgg * (2.25f * __builtin_fmaf(b * ia, ia, __builtin_fmaf(a * ib, ib, 3.0f - ab)));
        w[1] = Lgg * (k.nQuarkNew * q2 * __builtin_fmaf(inv_ab, 0.16666666666666667f, -0.375f));
        w[2] = Lqg * (__builtin_fmaf(b, b, 1.0f) * __builtin_fmaf(ia, ia, 0.44444444444444444f * ib));
        {
            const float sigT = 0.44444444444444444f * __builtin_fmaf(b, b, 1.0f) * (ia * ia);
            const float sigU = 0.44444444444444444f * __builtin_fmaf(a, a, 1.0f) * (ib * ib);
            const float sigTU = -0.29629629629629630f * inv_ab;
            const float sigST = 0.29629629629629630f * (b * b) * ia;
            const float csame = 0.5f * (sigU + sigTU - sigT);
            w[3] = __builtin_fmaf(sigT, LQQ, __builtin_fmaf(csame, Lsame, sigST * Lopp));
        }
        w[4] = Lopp * (q2 * __builtin_fmaf(inv_ab, 0.59259259259259259f, -1.3333333333333333f));
        w[5] = Lopp * (k.nQuarkNew * 0.44444444444444444f * q2);
        // a slot mask leaves out is 0 and adds nothing (a uniform branch, no select: the empty
        // volatile asm keeps the compiler from if-converting it)
        // Pythia clamps a negative sigma to 0 and keeps a NaN.  The bit-exact form does so literally;
        // in FP32_HW the clamp is left out, being a no-op here: every factor of these six weights is
        // >= 0 (luminosities of clipped or exponential densities, the prefactor) and so is every
        // bracket -- sums of positive terms, or 1/(6ab) - 3/8 >= 0.29 and 16/(27ab) - 4/3 >= 1.03
        // since ab <= 1/4; for q q' -> q q' with csame < 0, Lsame <= LQQ gives w3 >= LQQ (sigT +
        // csame) = LQQ (sigT + sigU + sigTU) / 2, the identical-quark bracket, which is at least
        // 0.93 max(sigT, sigU) -- far beyond FP32 rounding -- so no weight is negative or -0.
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            ws[i] = 0.0f;
            if (mask & (1 << i)) {
                if (BR) __asm__ volatile("");
                float wi = common * w[i];
                if (!HW) wi = (wi > 0.0f || wi != wi) ? wi : 0.0f;   // negative -> 0, NaN kept
                ws[i] = wi;
                tot += wi;
            }
        }
        if (sHo) { *sHo = sH; *tHo = -(sH * a); *x1o = x1; *x2o = x2; }
    } else {
        // Sigma2gg2QQbar / Sigma2qqbar2QQbar::sigmaKin with tHQ = -sH a, uHQ = -sH b, mu = m^2 / sH:
        //   gg:  [b/a - 9/4 b^2 + mu/a^2 (a (9/2 b - 1/2 + mu) - 4 mu)  +  (a <-> b)] / 6
        //   qqb: 4/9 (a^2 + b^2 + 2 mu)
        const float mu = k.fam[F].mu * r;
        const float cT = __builtin_fmaf(a, __builtin_fmaf(4.5f, b, -0.5f) + mu, -4.0f * mu);
        const float cU = __builtin_fmaf(b, __builtin_fmaf(4.5f, a, -0.5f) + mu, -4.0f * mu);
        const float sTS = __builtin_fmaf(b, ia, __builtin_fmaf(-2.25f * b, b, (mu * (ia * ia)) * cT));
        const float sUS = __builtin_fmaf(a, ib, __builtin_fmaf(-2.25f * a, a, (mu * (ib * ib)) * cU));
        float w[2];
        w[0] = Lgg * (0.16666666666666667f * (sTS + sUS));
        w[1] = Lopp * (0.44444444444444444f * __builtin_fmaf(2.0f, mu, q2));
        constexpr int lo = 4 + 2 * FAM;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            ws[lo + i] = 0.0f;
            if (mask & (1 << (lo + i))) {
                float wi = common * w[i];
                wi = (wi > 0.0f || wi != wi) ? wi : 0.0f;      // negative -> 0, NaN kept
                ws[lo + i] = wi;
                tot += wi;
            }
        }
        if (sHo) { *sHo = sH; *tHo = __builtin_fmaf(-sH, a, k.fam[F].m2); *x1o = x1; *x2o = x2; }
    }
    if (!HW && aux) {
        aux->alpS = alpS; aux->cf = cfo; aux->lx1 = lx1; aux->lx2 = lx2; aux->l1a = l1a; aux->l1b = l1b;
        aux->g1 = g1; aux->g2 = g2; aux->Q1 = Q1; aux->Q2 = Q2s;
    }
    return tot;
}

// The point: every family selected by mask, at the same uniforms.  HEAVY = false evaluates the
// massless family only (codes 111..116) and writes ws[6..9] = 0; kinematics outputs are then
// single values.  HEAVY = true: sHo .. x2o (if given) are arrays of 3, one entry per family,
// written for the families that mask selects.
template <bool HEAVY = false, bool HW = false, bool U24 = false>
__device__ __forceinline__ float had_point_f32(const HadConst<float>& k, int mask,
                                               float u0, float om0, float u1, float om1,
                                               float u2, float om2, float* ws, float* sHo, float* tHo,
                                               float* x1o, float* x2o, PointAux<float>* aux,
                                               const GridPdf<float>* gp, bool grid)
{
    if (!HEAVY) {
#pragma unroll
        for (int i = kLightProc; i < kMaxProc; ++i) ws[i] = 0.0f;
        return had_family_f32<0, HW, U24>(k, mask, u0, om0, u1, om1, u2, om2, ws, sHo, tHo, x1o, x2o, aux, gp, grid);
    }
#pragma unroll
    for (int i = 0; i < kMaxProc; ++i) ws[i] = 0.0f;
    float tot = 0.0f;
    if (mask & kFamMask[0])
        tot += had_family_f32<0, HW, U24, false>(k, mask, u0, om0, u1, om1, u2, om2, ws, sHo, tHo, x1o, x2o, aux, gp, grid);
    if (mask & kFamMask[1])
        tot += had_family_f32<1, HW, U24>(k, mask, u0, om0, u1, om1, u2, om2, ws, sHo ? sHo + 1 : nullptr, tHo ? tHo + 1 : nullptr,
                                 x1o ? x1o + 1 : nullptr, x2o ? x2o + 1 : nullptr, nullptr, gp, grid);
    if (mask & kFamMask[2])
        tot += had_family_f32<2, HW, U24>(k, mask, u0, om0, u1, om1, u2, om2, ws, sHo ? sHo + 2 : nullptr, tHo ? tHo + 2 : nullptr,
                                 x1o ? x1o + 2 : nullptr, x2o ? x2o + 2 : nullptr, nullptr, gp, grid);
    return tot;
}

// ------------------------------------------------------------------ FP64 hadronic point
// Same map and algebra in FP64 with the device libm (exp, log, expm1, sqrt); the per-point
// parity target against the host Pythia-form oracle is 1e-12 relative.
// alphaS of Pythia's order-0/1 AlphaStrong at Q2 (FP64); cf = 12 pi / b0 of the active flavours.
__host__ __device__ inline double alphas_f64(const HadConst<double>& k, double Q2, double* cfo = nullptr)
{
    if (k.as_order == 0.0) { if (cfo) *cfo = 0.0; return k.as_fixed; }
    if (Q2 < k.scale2Min) Q2 = k.scale2Min;
    const double lq = log(Q2);
    double lam, cf;
    if (Q2 > k.mt2 && k.nf6 != 0.0) { lam = k.lnL6s; cf = k.c6; }
    else if (Q2 > k.mb2)            { lam = k.lnL5s; cf = k.c5; }
    else if (Q2 > k.mc2)            { lam = k.lnL4s; cf = k.c4; }
    else                            { lam = k.lnL3s; cf = k.c3; }
    if (cfo) *cfo = cf;
    return cf / (lq - lam);
}

// Luminosities of the toy PDFs at (x1, x2) = (e^lx1, e^lx2) (FP64); l1a, l1b = ln(1 - x).
__host__ __device__ inline Lumi<double> toy_lumi_f64(const HadConst<double>& k, double lx1, double lx2, double& l1a,
                                                     double& l1b)
{
    Lumi<double> L;
    l1a = log(-expm1(lx1)); l1b = log(-expm1(lx2));
    const double g1 = exp(fma(k.bg, l1a, fma(k.ag, lx1, k.lnAg)));
    const double g2 = exp(fma(k.bg, l1b, fma(k.ag, lx2, k.lnAg)));
    const double S1 = exp(fma(k.bs, l1a, fma(k.as, lx1, k.lnAs)));
    const double S2 = exp(fma(k.bs, l1b, fma(k.as, lx2, k.lnAs)));
    const double V1 = exp(fma(k.bv, l1a, fma(k.av, lx1, k.lnAuv)));
    const double V2 = exp(fma(k.bv, l1b, fma(k.av, lx2, k.lnAuv)));
    const double Vd1 = k.rdu * V1, Vd2 = k.rdu * V2;
    const double U1 = V1 + S1, U2 = V2 + S2, D1 = Vd1 + S1, D2 = Vd2 + S2;
    const double Q1 = fma(k.nSea, S1, V1 + Vd1), Q2s = fma(k.nSea, S2, V2 + Vd2);
    const double SS = S1 * S2;
    L.g1 = g1; L.g2 = g2; L.Q1 = Q1; L.Q2 = Q2s;
    L.gg = g1 * g2;
    L.qg = fma(Q1, g2, g1 * Q2s);
    L.QQ = Q1 * Q2s;
    L.same = fma(U1, U2, fma(D1, D2, k.nSeaM2 * SS));
    L.opp = fma(U1 + D1, S2, fma(S1, U2 + D2, k.nSeaM4 * SS));
    return L;
}

// Bracket times luminosity of each process of a family (FP64): FAM 0 fills w[0..5] (codes
// 111..116), FAM 1, 2 fill w[0..1] (the family's gg and q qbar codes); a = (m^2 - tH)/sH,
// b = (m^2 - uH)/sH, mu = m^2 / sH (0 for FAM 0).  d sigma / dt = pi alpS^2 / sH^2 * bracket.
template <int FAM>
__host__ __device__ inline void brackets_f64(const HadConst<double>& k, double a, double b, double mu,
                                             const Lumi<double>& L, double* w)
{
    const double ab = a * b;
    const double inv_ab = 1.0 / ab;
    const double ia = b * inv_ab, ib = a * inv_ab;
    const double q2 = fma(a, a, b * b);
    if (FAM == 0) {
        w[0] = L.gg * (2.25 * fma(b * ia, ia, fma(a * ib, ib, 3.0 - ab)));
        w[1] = L.gg * (k.nQuarkNew * q2 * fma(inv_ab, 1.0 / 6.0, -0.375));
        w[2] = L.qg * (fma(b, b, 1.0) * fma(ia, ia, (4.0 / 9.0) * ib));
        {
            const double sigT = (4.0 / 9.0) * fma(b, b, 1.0) * (ia * ia);
            const double sigU = (4.0 / 9.0) * fma(a, a, 1.0) * (ib * ib);
            const double sigTU = -(8.0 / 27.0) * inv_ab;
            const double sigST = (8.0 / 27.0) * (b * b) * ia;
            const double csame = 0.5 * (sigU + sigTU - sigT);
            w[3] = fma(sigT, L.QQ, fma(csame, L.same, sigST * L.opp));
        }
        w[4] = L.opp * (q2 * fma(inv_ab, 16.0 / 27.0, -4.0 / 3.0));
        w[5] = L.opp * (k.nQuarkNew * (4.0 / 9.0) * q2);
    } else {
        const double cT = fma(a, fma(4.5, b, -0.5) + mu, -4.0 * mu);
        const double cU = fma(b, fma(4.5, a, -0.5) + mu, -4.0 * mu);
        const double sTS = fma(b, ia, fma(-2.25 * b, b, (mu * (ia * ia)) * cT));
        const double sUS = fma(a, ib, fma(-2.25 * a, a, (mu * (ib * ib)) * cU));
        w[0] = L.gg * ((sTS + sUS) / 6.0);
        w[1] = L.opp * ((4.0 / 9.0) * fma(2.0, mu, q2));
    }
}

template <int FAM>
__device__ __forceinline__ double had_family_f64(const HadConst<double>& k, int mask,
                                                 double u0, double om0, double u1, double om1,
                                                 double u2, double om2, double* ws, double* sHo,
                                                 double* tHo, double* x1o, double* x2o,
                                                 PointAux<double>* aux, const GridPdf<double>* gp, bool grid)
{
    const double LT = FAM == 0 ? k.LT : k.fam[FAM - 1].LT;
    const double targ = -(u0 * LT);
    const double r = exp(targ);
    const double yMax = 0.5 * (LT * om0);
    const double twoYMax = yMax + yMax;
    const double lx1 = -(twoYMax * om1), lx2 = -(twoYMax * u1);
    const double x1 = exp(lx1), x2 = exp(lx2);
    const double zMax = sqrt(-expm1(targ));
    const double omz = r / (1.0 + zMax);
    const double twoz = zMax + zMax;
    const double A = fma(twoz, om2, omz);
    const double B = fma(twoz, u2, omz);
    const double a = 0.5 * A, b = 0.5 * B;
    const double sH = (FAM == 0 ? k.fourPT2 : k.fam[FAM - 1].fourMT2) / r;
    const double pT2 = sH * (a * b);

    double cfo = 0.0;
    const double alpS = alphas_f64(k, k.xiR2 * pT2, &cfo);

    Lumi<double> L;
    double l1a = 0.0, l1b = 0.0;
    if (grid) {
        const double Q2F = gp->factFac2 * pT2;
        const double q = sqrt(Q2F);
        L = grid_lumi<double>(*gp, (int)(0.5 * k.nSea), x1, lx1, x2, lx2, q, log(q));
    } else {
        L = toy_lumi_f64(k, lx1, lx2, l1a, l1b);
    }

    const double common = (FAM == 0 ? k.pref : k.fam[FAM - 1].pref) * (alpS * alpS) * (LT * om0) * zMax * r;
    double tot = 0.0;
    constexpr int NW = FAM == 0 ? 6 : 2, lo = FAM == 0 ? 0 : 4 + 2 * FAM;
    double w[NW];
    brackets_f64<FAM>(k, a, b, FAM == 0 ? 0.0 : k.fam[FAM - 1].mu * r, L, w);
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        double wi = (mask & (1 << (lo + i))) ? common * w[i] : 0.0;
        wi = (wi > 0.0 || wi != wi) ? wi : 0.0;                // negative -> 0, NaN kept
        ws[lo + i] = wi;
        tot += wi;
    }
    if (sHo) {
        *sHo = sH; *tHo = FAM == 0 ? -(sH * a) : fma(-sH, a, k.fam[FAM - 1].m2); *x1o = x1; *x2o = x2;
    }
    if (aux) {
        aux->alpS = alpS; aux->cf = cfo; aux->lx1 = lx1; aux->lx2 = lx2; aux->l1a = l1a; aux->l1b = l1b;
        aux->g1 = L.g1; aux->g2 = L.g2; aux->Q1 = L.Q1; aux->Q2 = L.Q2;
    }
    return tot;
}

template <bool HEAVY = false>
__device__ __forceinline__ double had_point_f64(const HadConst<double>& k, int mask,
                                                double u0, double om0, double u1, double om1,
                                                double u2, double om2, double* ws, double* sHo,
                                                double* tHo, double* x1o, double* x2o,
                                                PointAux<double>* aux, const GridPdf<double>* gp, bool grid)
{
    if (!HEAVY) {
#pragma unroll
        for (int i = kLightProc; i < kMaxProc; ++i) ws[i] = 0.0;
        return had_family_f64<0>(k, mask, u0, om0, u1, om1, u2, om2, ws, sHo, tHo, x1o, x2o, aux, gp, grid);
    }
#pragma unroll
    for (int i = 0; i < kMaxProc; ++i) ws[i] = 0.0;
    double tot = 0.0;
    if (mask & kFamMask[0])
        tot += had_family_f64<0>(k, mask, u0, om0, u1, om1, u2, om2, ws, sHo, tHo, x1o, x2o, aux, gp, grid);
    if (mask & kFamMask[1])
        tot += had_family_f64<1>(k, mask, u0, om0, u1, om1, u2, om2, ws, sHo ? sHo + 1 : nullptr, tHo ? tHo + 1 : nullptr,
                                 x1o ? x1o + 1 : nullptr, x2o ? x2o + 1 : nullptr, nullptr, gp, grid);
    if (mask & kFamMask[2])
        tot += had_family_f64<2>(k, mask, u0, om0, u1, om1, u2, om2, ws, sHo ? sHo + 2 : nullptr, tHo ? tHo + 2 : nullptr,
                                 x1o ? x1o + 2 : nullptr, x2o ? x2o + 2 : nullptr, nullptr, gp, grid);
    return tot;
}

// ------------------------------------------------------------------ partonic gg -> gg, FP64
// w(c) = CONV * pi alphaS^2 / s^2 * B_gg(a, b) * (2 cMax)(s / 2), a = (1 - c)/2, b = (1 + c)/2.
// pref = CONV pi alphaS^2 cMax / s.
__device__ __forceinline__ double c1_weight(double pref, double cMax, double u)
{
    const double c = cMax * (2.0 * u - 1.0);
    const double a = 0.5 * (1.0 - c), b = 0.5 * (1.0 + c);
    const double ab = a * b;
    const double inv_ab = 1.0 / ab;
    const double ia = b * inv_ab, ib = a * inv_ab;
    return pref * (2.25 * fma(b * ia, ia, fma(a * ib, ib, 3.0 - ab)));
}

}  // namespace mi
