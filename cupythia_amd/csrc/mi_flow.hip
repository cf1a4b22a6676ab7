// mi_flow.hip -- spline-coupling normalizing flow layered on the VEGAS map: forward
// (normalizing, y -> z, log|det dz/dy|), inverse (sampling, z -> y), backward (parameter
// gradients of sum_i g_i log q(y_i)), variance-loss training with Adam, and the flow sampler of
// the integration -- all on the device.
//
// Architecture: 3-D, 6 rational-quadratic spline couplings on [0,1] (Durkan, Bekasov, Murray,
// Papamakarios, "Neural Spline Flows", NeurIPS 2019) with 16 bins; layer l transforms
// coordinate l % 3 conditioned on the other two through an MLP 2 -> 64 -> 64 -> 49 (ReLU):
// 16 width and 16 height logits (softmax, minimum 1e-3) and 17 knot derivatives
// (1e-3 + softplus).  Weights are stored as BF16 and the MLP runs on the matrix cores
// (v_mfma_f32_32x32x16_bf16, FP32 accumulation); the splines run on VALU in FP32.
//
// Execution: a wave owns 64 points per tile and walks tiles with a grid stride.  The forward /
// inverse kernel runs all six layers of a tile in one launch with every layer's weights in LDS
// (111 KB) and the activations in registers; the backward kernel runs one layer per launch from
// the stored layer inputs, with that layer's weights (17.6 KB) and the activations in LDS.
//
// Training (i-flow / neural importance sampling with the variance loss): points y = T^-1(z)
// from the current flow, w = f(V(y)) J_V(y) |dy/dz|; the gradient of E_q[w^2] is
// -E_q[w^2 grad log q(y)] (score form, samples held fixed), so the loss surrogate is
// sum_i g_i log q(y_i) with g_i = -w_i^2 / (N <w>^2).  The initial flow is the identity (last
// layers zero), so flow + VEGAS starts as VEGAS; the weight f / q stays the unbiased estimator.
#include "mi_flow_common.h"

#include <cmath>
#include <cstring>
#include <algorithm>
#include <vector>

namespace mi {

// ------------------------------------------------------------------------------------ forward / inverse
// The forward / inverse kernel runs all FL layers of a 64-point tile (lane = point) with the
// coordinates in registers, and keeps the MLP in the matrix-core registers as well: it computes
// C = W H^T (hidden units in rows, points in columns), after which a lane holds 16 hidden units of
// one point (units 32 mb + 8 i + 4 (l / 32) + r in register 4 i + r of block (mb, nb)).  Those
// registers are the next GEMM's B operand as they stand once the contraction index is permuted:
// slot k of a 16-wide K step carries unit swz(k) (bits 2 and 3 of k exchanged), and the weights
// are stored with their columns permuted the same way.  The 2 -> 64 layer is a K = 16 step too:
// the two conditioners enter split into three BF16 parts each (exact), the bias b1 likewise
// against a constant-one operand.  The layer-3 outputs are brought to the point's own lane with
// v_permlane32_swap, and the spline runs on them in registers.

// MLP outputs o[0 .. FO-1] of the point of this lane for conditioners (c0, c1).
__device__ __forceinline__ void chain_mlp(const FwdLayer &W, float c0, float c1, float (&o)[FO])
{
    const int lane = threadIdx.x & 63, hh = lane >> 5, r32 = lane & 31;
    // ---- layer 1: K slots (c0, c1) x (hi, mid, lo) against (W1[m][0], W1[m][1]) x 3, and
    // (1, 1, 1) against the three parts of b1
    const CondFrag cf = cond_frag(c0, c1);
    f32x16 C[2][2];
    const f32x16 zero = {};
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        const bf16x8 a = as_frag(W.W1f[hh][32 * mb + r32]);
        C[mb][0] = mfma16(a, as_frag(cf.frag[0]), zero);
        C[mb][1] = mfma16(a, as_frag(cf.frag[1]), zero);
    }
    // ---- layers 2 and 3
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
        bf16x8 B[4][2];
#pragma unroll
        for (int kk = 0; kk < 4; ++kk)
#pragma unroll
            for (int nb = 0; nb < 2; ++nb) B[kk][nb] = relu_frag(C[kk >> 1][nb], 8 * (kk & 1));
        const uint4(*Wf)[2][64] = layer == 0 ? W.W2f : W.W3f;
        const float *bias = layer == 0 ? W.b2 : W.b3;
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            C[mb][0] = bias_block(bias, mb, hh);
            C[mb][1] = C[mb][0];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const bf16x8 a = as_frag(Wf[kk][hh][32 * mb + r32]);
                C[mb][0] = mfma16(a, B[kk][0], C[mb][0]);
                C[mb][1] = mfma16(a, B[kk][1], C[mb][1]);
            }
        }
    }
    rows_to_lane(C, o);
}

// Spline of one point from its MLP outputs (widths o[0..15], heights o[16..31], knot-derivative
// logits o[32..48]): the forward map x -> y with +log S'(x), or the inverse y -> x with
// -log S'(x).  Knot j of a side is kMin j + (1 - FK kMin) C_j / C_FK with C_j the cumulative sum
// of the softmax numerators e^(o - max); the bin is located on the input side and its six
// quantities are carried along the scan by selects.
template <bool INVERSE>
__device__ __forceinline__ float chain_spline(const float (&o)[FO], float v, float &lj)
{
    constexpr float L2E = 1.4426950408889634f;
    float mw = o[0], mh = o[FK];
#pragma unroll
    for (int j = 1; j < FK; ++j) { mw = fmaxf(mw, o[j]); mh = fmaxf(mh, o[FK + j]); }
    float ew[FK], eh[FK], Cw[FK + 1], Ch[FK + 1];
    const float nw = -mw * L2E, nh = -mh * L2E;
    Cw[0] = 0.0f; Ch[0] = 0.0f;
#pragma unroll
    for (int j = 0; j < FK; ++j) {
        ew[j] = exp2_hw(__builtin_fmaf(o[j], L2E, nw));
        eh[j] = exp2_hw(__builtin_fmaf(o[FK + j], L2E, nh));
        Cw[j + 1] = Cw[j] + ew[j];
        Ch[j + 1] = Ch[j] + eh[j];
    }
    const float rw = (1.0f - FK * kMinW) * rcp_hw(Cw[FK]), rh = (1.0f - FK * kMinH) * rcp_hw(Ch[FK]);
    // S: the side the input lives on (heights for the inverse), T: the other side
    const float *eS = INVERSE ? eh : ew, *CS = INVERSE ? Ch : Cw, *eT = INVERSE ? ew : eh, *CT = INVERSE ? Cw : Ch;
    const float rS = INVERSE ? rh : rw, rT = INVERSE ? rw : rh;
    const float mS = INVERSE ? kMinH : kMinW, mT = INVERSE ? kMinW : kMinH;
    float sK = 0.0f, seS = eS[0], sCT = 0.0f, seT = eT[0], sk = 0.0f, sd0 = o[2 * FK], sd1 = o[2 * FK + 1];
#pragma unroll
    for (int j = 1; j < FK; ++j) {
        const float K = __builtin_fmaf(rS, CS[j], mS * (float)j);
        const bool c = v >= K;
        sK = c ? K : sK;
        seS = c ? eS[j] : seS;
        sCT = c ? CT[j] : sCT;
        seT = c ? eT[j] : seT;
        sk = c ? (float)j : sk;
        sd0 = c ? o[2 * FK + j] : sd0;
        sd1 = c ? o[2 * FK + j + 1] : sd1;
    }
    const float sS = __builtin_fmaf(rS, seS, mS);                         // bin size, input side
    const float tK = __builtin_fmaf(rT, sCT, mT * sk), sT = __builtin_fmaf(rT, seT, mT);
    const float dk = kMinD + softplus_hw(sd0), dk1 = kMinD + softplus_hw(sd1);
    if (!INVERSE) {
        const float xk = sK, wk = sS, yk = tK, hk = sT;
        const float iw = rcp_hw(wk);
        const float xi = (v - xk) * iw, s = hk * iw, t = xi * (1.0f - xi);
        const float den = __builtin_fmaf(dk1 + dk - 2.0f * s, t, s);
        const float iden = rcp_hw(den);
        const float y = __builtin_fmaf(hk * __builtin_fmaf(s * xi, xi, dk * t), iden, yk);
        const float omx = 1.0f - xi;
        const float n2 = __builtin_fmaf(dk1 * xi, xi, __builtin_fmaf(2.0f * s, t, dk * omx * omx));
        const float sd = s * iden;
        lj = log_hw(sd * sd * n2);
        return y;
    } else {
        const float yk = sK, hk = sS, xk = tK, wk = sT;
        const float s = hk * rcp_hw(wk);
        const float dy = v - yk;
        const float A = dk1 + dk - 2.0f * s;
        const float qa = __builtin_fmaf(hk, s - dk, dy * A);
        const float qb = __builtin_fmaf(hk, dk, -dy * A);
        const float qc = -s * dy;
        const float disc = fmaxf(__builtin_fmaf(qb, qb, -4.0f * qa * qc), 0.0f);
        float xi = (2.0f * qc) * rcp_hw(-qb - __builtin_amdgcn_sqrtf(disc));
        xi = fminf(fmaxf(xi, 0.0f), 1.0f);
        const float t = xi * (1.0f - xi), omx = 1.0f - xi;
        const float den = __builtin_fmaf(A, t, s);
        const float n2 = __builtin_fmaf(dk1 * xi, xi, __builtin_fmaf(2.0f * s, t, dk * omx * omx));
        const float sd = s * rcp_hw(den);
        lj = -log_hw(sd * sd * n2);
        return __builtin_fmaf(xi, wk, xk);
    }
}

// All FL layers over n points, normalizing (x^(0) -> x^(FL)) or sampling (x^(FL) -> x^(0)).
// Input: xin (n x 3), or with SAMPLE the three 24-bit uniforms of Philox stream `sid` of events
// ctr0 + i.  Output: xout (n x 3, optional, may alias xin), ld (n) = the summed log-Jacobian
// (written, not accumulated).  keep (optional): every layer input of the normalizing direction,
// x^(j) at keep + j * kstride (x^(0) is the caller's input in the normalizing direction and is
// not rewritten).  One block per CU holds all layers' weights in LDS (111 KB); blocks have FWV
// waves for large runs, else 8, or 4 (one per SIMD) when there are too few tiles (fwd_waves).
template <bool INVERSE, bool SAMPLE>
__global__ __launch_bounds__(64 * FWV) void k_flow_chain(const FwdLayer *__restrict__ layers, uint64_t n,
                                                         const float *xin, float *xout,
                                                         float *__restrict__ ld, float *__restrict__ keep,
                                                         uint64_t kstride, uint64_t key, uint64_t ctr0, uint32_t sid)
{
    __shared__ FwdLayer W[FL];
    {
        // eight 16-byte loads of a work-item in flight before their LDS writes (a load-wait-write
        // loop cost ~28 memory round trips per work-item at 4 waves per block)
        constexpr int NV = (int)(sizeof(W) / 16);
        const uint4 *src = reinterpret_cast<const uint4 *>(layers);
        uint4 *dst = reinterpret_cast<uint4 *>(W);
        const int nt = (int)blockDim.x;
        for (int j0 = threadIdx.x; j0 < NV; j0 += 8 * nt) {
            const auto ld = [=](int k) { return j0 + k * nt < NV ? src[j0 + k * nt] : make_uint4(0u, 0u, 0u, 0u); };
            const auto st = [=](int k, uint4 v) { if (j0 + k * nt < NV) dst[j0 + k * nt] = v; };
            const uint4 v0 = ld(0), v1 = ld(1), v2 = ld(2), v3 = ld(3), v4 = ld(4), v5 = ld(5), v6 = ld(6), v7 = ld(7);
            st(0, v0); st(1, v1); st(2, v2); st(3, v3); st(4, v4); st(5, v5); st(6, v6); st(7, v7);
        }
        __syncthreads();
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t ntiles = (n + 63) / 64;
    const int nwave = blockDim.x >> 6;
    for (uint64_t tile = (uint64_t)blockIdx.x * nwave + wave; tile < ntiles; tile += (uint64_t)gridDim.x * nwave) {
        const uint64_t i = tile * 64 + lane;
        const bool live = i < n;
        float x[3] = {0.5f, 0.5f, 0.5f};
        if (SAMPLE) {
            const U4 r = philox_event(key, ctr0 + i, sid);
            x[0] = (float)k24(r.x) * (1.0f / 16777216.0f);
            x[1] = (float)k24(r.y) * (1.0f / 16777216.0f);
            x[2] = (float)k24(r.z) * (1.0f / 16777216.0f);
        } else if (live) {
            x[0] = xin[3 * i]; x[1] = xin[3 * i + 1]; x[2] = xin[3 * i + 2];
        }
        if (INVERSE && keep && live) {
            float *k = keep + (size_t)FL * kstride + 3 * i;
            k[0] = x[0]; k[1] = x[1]; k[2] = x[2];
        }
        float lsum = 0.0f;
#pragma unroll 1
        for (int s = 0; s < FL; ++s) {
            int l = INVERSE ? FL - 1 - s : s;
            // l opaque (an SGPR the compiler cannot rewrite in s): otherwise the inverse direction's
            // W[FL - 1 - s] becomes a 95 KB constant offset past the ds_read offset field, and every
            // weight fragment read costs two address additions
            __asm__ volatile("" : "+s"(l));
            const int d = l % 3;
            float o[FO];
            chain_mlp(W[l], pick3(x, cond0(d)), pick3(x, cond1(d)), o);
            float lj;
            set3(x, d, chain_spline<INVERSE>(o, pick3(x, d), lj));
            lsum += lj;
            if (keep && live) {
                float *k = keep + (size_t)(INVERSE ? l : l + 1) * kstride + 3 * i;
                k[0] = x[0]; k[1] = x[1]; k[2] = x[2];
            }
        }
        if (live) {
            if (xout) { xout[3 * i] = x[0]; xout[3 * i + 1] = x[1]; xout[3 * i + 2] = x[2]; }
            ld[i] = lsum;
        }
    }
}

// ------------------------------------------------------------------------------------ small kernels
// The weight images of layer blockIdx.x from the FP32 parameters (BF16, round to nearest even).
constexpr int kPackSplit = 8;   // blocks per layer of the pack kernel (blockIdx.y)
__global__ void k_flow_pack(const float *__restrict__ params, FwdLayer *__restrict__ fwd, BwdExtra *__restrict__ bwd)
{
    const int l = blockIdx.x;
    const int t0 = blockIdx.y * blockDim.x + threadIdx.x, ts = gridDim.y * blockDim.x;
    const float *p = params + (size_t)l * FPL;
    FwdLayer &f = fwd[l];
    BwdExtra &e = bwd[l];
    for (int m = t0; m < FH; m += ts) {
        const uint32_t w = (uint32_t)f2bf(p[PW1 + 2 * m]) | ((uint32_t)f2bf(p[PW1 + 2 * m + 1]) << 16);
        f.W1f[0][m] = make_uint4(w, w, w, 0u);
        const float b = p[PB1 + m];
        const uint16_t hi = f2bf(b);
        const float r = b - bf2f(hi);
        const uint16_t mid = f2bf(r);
        const uint16_t lo = f2bf(r - bf2f(mid));
        f.W1f[1][m] = make_uint4((uint32_t)hi | ((uint32_t)mid << 16), (uint32_t)lo, 0u, 0u);
        f.b2[m] = p[PB2 + m];
        f.b3[m] = m < FO ? p[PB3 + m] : 0.0f;
    }
    for (int idx = t0; idx < 4 * 2 * 64; idx += ts) {
        const int kk = idx >> 7, hh = (idx >> 6) & 1, m = idx & 63;
        uint32_t w2[4], w3[4];
        for (int j = 0; j < 8; j += 2) {
            const int ka = swz(16 * kk + 8 * hh + j), kb = swz(16 * kk + 8 * hh + j + 1);
            w2[j / 2] = (uint32_t)f2bf(p[PW2 + m * FH + ka]) | ((uint32_t)f2bf(p[PW2 + m * FH + kb]) << 16);
            w3[j / 2] = m < FO ? (uint32_t)f2bf(p[PW3 + m * FH + ka]) | ((uint32_t)f2bf(p[PW3 + m * FH + kb]) << 16) : 0u;
        }
        f.W2f[kk][hh][m] = make_uint4(w2[0], w2[1], w2[2], w2[3]);
        f.W3f[kk][hh][m] = make_uint4(w3[0], w3[1], w3[2], w3[3]);
        // transposed: column m of W2 / W3 with its rows permuted
        uint32_t t2[4], t3[4];
        for (int j = 0; j < 8; j += 2) {
            const int ka = swz(16 * kk + 8 * hh + j), kb = swz(16 * kk + 8 * hh + j + 1);
            t2[j / 2] = (uint32_t)f2bf(p[PW2 + ka * FH + m]) | ((uint32_t)f2bf(p[PW2 + kb * FH + m]) << 16);
            const uint32_t lo = ka < FO ? (uint32_t)f2bf(p[PW3 + ka * FH + m]) : 0u;
            const uint32_t hi = kb < FO ? (uint32_t)f2bf(p[PW3 + kb * FH + m]) : 0u;
            t3[j / 2] = lo | (hi << 16);
        }
        e.W2Tf[kk][hh][m] = make_uint4(t2[0], t2[1], t2[2], t2[3]);
        e.W3Tf[kk][hh][m] = make_uint4(t3[0], t3[1], t3[2], t3[3]);
    }
    for (int idx = t0; idx < 4 * 2 * 2; idx += ts) {
        const int kk = idx >> 2, hh = (idx >> 1) & 1, c = idx & 1;
        uint32_t t1[4];
        for (int j = 0; j < 8; j += 2) {
            const int ka = swz(16 * kk + 8 * hh + j), kb = swz(16 * kk + 8 * hh + j + 1);
            t1[j / 2] = (uint32_t)f2bf(p[PW1 + 2 * ka + c]) | ((uint32_t)f2bf(p[PW1 + 2 * kb + c]) << 16);
        }
        e.W1Tf[kk][hh][c] = make_uint4(t1[0], t1[1], t1[2], t1[3]);
    }
}

// Writes parameter j (0 .. FPL-1) of layer l, value v, into every slot of the weight images that
// carries it: the permutation k_flow_pack applies to a whole layer, one parameter at a time (the
// padding slots keep the zeros of the last full pack).
__device__ __forceinline__ uint16_t &img16(uint4 &q, int jj) { return reinterpret_cast<uint16_t *>(&q)[jj]; }
__device__ void pack_param(int l, int j, float v, FwdLayer *__restrict__ fwd, BwdExtra *__restrict__ bwd)
{
    FwdLayer &f = fwd[l];
    BwdExtra &e = bwd[l];
    if (j < PB1) {                                       // W1[m][c]
        const int m = j >> 1, c = j & 1, t = swz(m);
        const uint16_t b = f2bf(v);
        img16(f.W1f[0][m], c) = b;
        img16(f.W1f[0][m], 2 + c) = b;
        img16(f.W1f[0][m], 4 + c) = b;
        img16(e.W1Tf[t >> 4][(t >> 3) & 1][c], t & 7) = b;
    } else if (j < PW2) {                                // b1[m] in three parts
        const int m = j - PB1;
        const uint16_t hi = f2bf(v);
        const float r = v - bf2f(hi);
        const uint16_t mid = f2bf(r);
        img16(f.W1f[1][m], 0) = hi;
        img16(f.W1f[1][m], 1) = mid;
        img16(f.W1f[1][m], 2) = f2bf(r - bf2f(mid));
    } else if (j < PB2) {                                // W2[m][k]
        const int m = (j - PW2) >> 6, k = (j - PW2) & 63, sk = swz(k), sm = swz(m);
        const uint16_t b = f2bf(v);
        img16(f.W2f[sk >> 4][(sk >> 3) & 1][m], sk & 7) = b;
        img16(e.W2Tf[sm >> 4][(sm >> 3) & 1][k], sm & 7) = b;
    } else if (j < PW3) {
        f.b2[j - PB2] = v;
    } else if (j < PB3) {                                // W3[o][u]
        const int o = (j - PW3) >> 6, u = (j - PW3) & 63, su = swz(u), so = swz(o);
        const uint16_t b = f2bf(v);
        img16(f.W3f[su >> 4][(su >> 3) & 1][o], su & 7) = b;
        img16(e.W3Tf[so >> 4][(so >> 3) & 1][u], so & 7) = b;
    } else {
        f.b3[j - PB3] = v;
    }
}

// Weight of a flow point: y -> VEGAS map -> FP32 integrand, times J_V and |dy/dz| = exp(ld).
// Writes w (optional) and one kRec record per block.  HEAVY: mask may select codes 121..124;
// HW: the integrand with the hardware transcendentals (MI_PREC_FP32_HW); STATS: only the record's
// sum w and sum w^2 (the training step's statistics; the other slots are not written).  The sums
// per process slot run as in the integration kernel: FP64 sums, the FP32 maximum as the integer
// order of its bits, slots the mask leaves out skipped by a uniform branch; the next point's y
// and ld are loaded before this point's integrand, so their latency hides behind it.  GRID: the
// grid PDFs (a template argument, so that the toy-PDF kernels do not carry the grid's registers).
template <bool HEAVY, bool HW, bool STATS, bool GRID>
__global__ __launch_bounds__(kBlock) void k_flow_weight(HadConst<float> k, int mask, const float *__restrict__ gtab,
                                                        int nbins, uint64_t n, const float *__restrict__ y,
                                                        const float *__restrict__ ld, float *__restrict__ wout,
                                                        double *__restrict__ partials, GridPdf<float> gpdf)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float *stab = reinterpret_cast<float *>(smem);
    for (int i = threadIdx.x; i < 9 * nbins; i += blockDim.x) stab[i] = gtab[i];
    __syncthreads();
    constexpr int NP = STATS ? 0 : (HEAVY ? kMaxProc : kLightProc);
    uint32_t mx[NP + 1];
    double sw[NP + 1], sw2[NP + 1];
#pragma unroll
    for (int p = 0; p <= NP; ++p) { mx[p] = 0u; sw[p] = 0.0; sw2[p] = 0.0; }
    uint32_t nnz = 0;
    const float nbf = (float)nbins;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    float ny[3] = {0.0f, 0.0f, 0.0f}, nld = 0.0f;
    if (i < n) { ny[0] = y[3 * i]; ny[1] = y[3 * i + 1]; ny[2] = y[3 * i + 2]; nld = ld[i]; }
    for (; i < n; i += stride) {
        const float yv[3] = {ny[0], ny[1], ny[2]}, lj = nld;
        const uint64_t inext = i + stride;
        if (inext < n) { ny[0] = y[3 * inext]; ny[1] = y[3 * inext + 1]; ny[2] = y[3 * inext + 2]; nld = ld[inext]; }
        float u[3], om[3], jac = 1.0f;
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) {
            const float yy = fminf(fmaxf(yv[dd], 0.0f), 1.0f);
            const float t = yy * nbf;
            const int b = min((int)t, nbins - 1);
            const float frac = t - (float)b;
            const float *q = stab + 3 * (dd * nbins + b);
            u[dd] = __builtin_fmaf(frac, q[1], q[0]);
            om[dd] = __builtin_fmaf(-frac, q[1], q[2]);
            jac = jac * (nbf * q[1]);
        }
        float ws[kMaxProc];
        const float f = had_point_f32<HEAVY, HW>(k, mask, u[0], om[0], u[1], om[1], u[2], om[2], ws, nullptr, nullptr,
                                                 nullptr, nullptr, nullptr, &gpdf, GRID);
        const float ej = expf(lj);
        const float wt = f * jac * ej;
        if (wout) wout[i] = wt;
        const double dw = (double)wt;
        sw[0] += dw;
        sw2[0] = fma(dw, dw, sw2[0]);
        if (STATS) continue;
        mx[0] = run_max(mx[0], wt);
        nnz += (wt > 0.0f) ? 1u : 0u;
        const float scale = jac * ej;
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            if (!((mask >> p) & 1)) continue;                   // uniform: the slot stays 0
            if (!HEAVY) __asm__ volatile("");                   // kept a branch, not selects
            const float wp = ws[p] * scale;
            const double dp = (double)wp;
            sw[1 + p] += dp;
            sw2[1 + p] = fma(dp, dp, sw2[1 + p]);
            mx[1 + p] = run_max(mx[1 + p], wp);
        }
    }
    if (STATS) {
        __shared__ double red[kBlock / 64][2];
        const double s = wave_sum(sw[0]), q = wave_sum(sw2[0]);
        if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6][0] = s; red[threadIdx.x >> 6][1] = q; }
        __syncthreads();
        if (threadIdx.x < 2) {
            double t = red[0][threadIdx.x];
            for (int w = 1; w < kBlock / 64; ++w) t += red[w][threadIdx.x];
            partials[(size_t)blockIdx.x * kRec + (threadIdx.x == 0 ? kRS : kRQ)] = t;
        }
    } else {
        double mxd[NP + 1];
#pragma unroll
        for (int p = 0; p <= NP; ++p) mxd[p] = max_value<float>(mx[p]);
        block_record_sums<kBlock, NP>(sw, sw2, mxd, (double)nnz, partials + (size_t)blockIdx.x * kRec);
    }
}

// Adam's step (Kingma, Ba 2015) with bias corrections c1 = 1 - b1^t, c2 = 1 - b2^t; the updated
// parameters also go into the weight images fwd, bwd.
struct AdamStep {
    float *p, *m, *v;
    float lr, b1, b2, eps, c1, c2;
    FwdLayer *fwd;
    BwdExtra *bwd;
};

// Parameter gradient of all layers: grad[l*FPL + j] = sum over r < rows of
// parts[(l*stride + r)*FPLP + j], in a fixed order (four interleaved row groups, then their sum);
// with adam.p set, also Adam's step of parameter l*FPL + j.
__global__ __launch_bounds__(256) void k_grad_sum(const float *__restrict__ parts, uint32_t rows, uint32_t stride,
                                                  float *__restrict__ grad, AdamStep adam)
{
    __shared__ float red[4][64];
    const uint32_t c = blockIdx.x * 64 + (threadIdx.x & 63), rg = threadIdx.x >> 6;
    float s = 0.0f;
    if (c < (uint32_t)(FL * FPL)) {
        const uint32_t layer = c / FPL, j = c % FPL;
        const float *q = parts + (size_t)layer * stride * FPLP + j;
#pragma unroll 4
        for (uint32_t r = rg; r < rows; r += 4) s += q[(size_t)r * FPLP];
    }
    red[rg][threadIdx.x & 63] = s;
    __syncthreads();
    if (threadIdx.x < 64 && c < (uint32_t)(FL * FPL)) {
        const float gi = (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
        grad[c] = gi;
        if (adam.p) {
            const float mi = adam.b1 * adam.m[c] + (1.0f - adam.b1) * gi;
            const float vi = adam.b2 * adam.v[c] + (1.0f - adam.b2) * gi * gi;
            adam.m[c] = mi;
            adam.v[c] = vi;
            const float pn = adam.p[c] - adam.lr * (mi / adam.c1) / (sqrtf(vi / adam.c2) + adam.eps);
            adam.p[c] = pn;
            pack_param((int)(c / FPL), (int)(c % FPL), pn, adam.fwd, adam.bwd);
        }
    }
}

// Adds a kRec record (sum slots add, max slots max) into an accumulating record.
__global__ void k_add_record(double *__restrict__ acc, const double *__restrict__ rec, int first)
{
    const int i = threadIdx.x;
    if (i >= kRec) return;
    if (first) { acc[i] = rec[i]; return; }
    acc[i] = rec_is_max(i) ? fmax(acc[i], rec[i]) : acc[i] + rec[i];
}

}  // namespace mi

// ============================================================================ host
using namespace mi;

namespace mi {

struct FlowState {
    float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    FwdLayer *fwd = nullptr;        // weight images of the kernels (k_flow_pack)
    BwdExtra *bwd = nullptr;
    // work buffers sized for `cap` points
    uint64_t cap = 0;
    float *xs = nullptr;        // (FL + 1) x cap x 3: layer inputs in the normalizing direction
    float *ld = nullptr, *w = nullptr, *abar = nullptr;
    float *gp = nullptr;        // per-block layer gradients, FL x gp_rows x FPLP
    uint32_t gp_rows = 0;
    double *hist = nullptr;     // training history (sum w, sum w^2 per step)
    int hist_cap = 0;
    double *acc = nullptr;      // accumulating record
    int step = 0;
    bool ready = false;
};

}  // namespace mi

namespace {

int ferr(mi_ctx *ctx, int code, const char *msg)
{
    ctx->err = msg;
    return code;
}

#define FCHK(call)                                                                                  \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            ctx->err = std::string(#call) + ": " + hipGetErrorString(e_);                          \
            return MI_ERR_HIP;                                                                      \
        }                                                                                           \
    } while (0)

// Host Philox4x32-10 (the same counter map as the device) for the initial weights.
void philox_host(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t o[4])
{
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    o[0] = c0; o[1] = c1; o[2] = c2; o[3] = c3;
}

int num_cu(mi_ctx *ctx) { return ctx->num_cu > 0 ? ctx->num_cu : 256; }

// Grids: the forward / inverse kernel (111 KB of LDS) and the backward kernel (159 KB) run one
// block per CU and walk the tiles with a grid stride.
// Waves per block of the forward / inverse kernel for n points, and its grid.
int fwd_waves(mi_ctx *ctx, uint64_t n)
{
    const uint64_t ntiles = (n + 63) / 64;
    // 12 waves (3 per SIMD) once every wave walks many tiles (3% faster at 2^26 points); 8 where
    // the grid stride would leave the waves uneven; 4 (one per SIMD) below two tiles per SIMD
    const uint64_t cu = (uint64_t)num_cu(ctx);
    if (ntiles >= cu * FWV * 16) return FWV;
    return ntiles < cu * 8 ? 4 : 8;
}
uint32_t fwd_blocks(mi_ctx *ctx, uint64_t n)
{
    const uint64_t ntiles = (n + 63) / 64, nw = (uint64_t)fwd_waves(ctx, n);
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((ntiles + nw - 1) / nw, (uint64_t)num_cu(ctx)));
}
uint32_t bwd_blocks(mi_ctx *ctx, uint64_t n)
{
    const uint64_t ntiles = (n + 63) / 64;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((ntiles + FWB - 1) / FWB, (uint64_t)num_cu(ctx)));
}

int flow_alloc(mi_ctx *ctx, uint64_t npts)
{
    FlowState &F = *ctx->flow;
    if (npts > F.cap) {
        (void)hipFree(F.xs); (void)hipFree(F.ld); (void)hipFree(F.w); (void)hipFree(F.abar);
        F.xs = nullptr; F.ld = F.w = F.abar = nullptr; F.cap = 0;
        FCHK(hipMalloc(&F.xs, sizeof(float) * 3 * (FL + 1) * npts));
        FCHK(hipMalloc(&F.ld, sizeof(float) * npts));
        FCHK(hipMalloc(&F.w, sizeof(float) * npts));
        FCHK(hipMalloc(&F.abar, sizeof(float) * 3 * npts));
        F.cap = npts;
    }
    const uint32_t rows = (uint32_t)num_cu(ctx);                  // one gradient row per backward block
    if (rows > F.gp_rows) {
        (void)hipFree(F.gp);
        F.gp = nullptr;
        F.gp_rows = 0;
        FCHK(hipMalloc(&F.gp, sizeof(float) * (size_t)FPLP * FL * rows));
        F.gp_rows = rows;
    }
    return MI_OK;
}

int check_flow(mi_ctx *ctx)
{
    if (!ctx) return MI_ERR_ARG;
    if (!ctx->flow || !ctx->flow->ready) return ferr(ctx, MI_ERR_STATE, "flow not initialised (mi_flow_init)");
    return MI_OK;
}

// All layers in one launch (k_flow_chain): normalizing (inverse = false) from xin, or sampling
// from xin or, with sample, from Philox stream sid of events ctr0 ..; keep = F.xs stores every
// layer input of the normalizing direction.
hipError_t run_chain(mi_ctx *ctx, bool inverse, bool sample, uint64_t n, const float *xin, float *xout, float *ld,
                     float *keep, uint64_t key = 0, uint64_t ctr0 = 0, uint32_t sid = 0)
{
    FlowState &F = *ctx->flow;
    const uint32_t blocks = fwd_blocks(ctx, n);
    const dim3 threads(64 * fwd_waves(ctx, n));
    const uint64_t ks = 3 * F.cap;
    if (inverse && sample)
        hipLaunchKernelGGL((k_flow_chain<true, true>), dim3(blocks), threads, 0, ctx->stream, F.fwd, n, xin, xout,
                           ld, keep, ks, key, ctr0, sid);
    else if (inverse)
        hipLaunchKernelGGL((k_flow_chain<true, false>), dim3(blocks), threads, 0, ctx->stream, F.fwd, n, xin,
                           xout, ld, keep, ks, key, ctr0, sid);
    else if (!sample)
        hipLaunchKernelGGL((k_flow_chain<false, false>), dim3(blocks), threads, 0, ctx->stream, F.fwd, n, xin,
                           xout, ld, keep, ks, key, ctr0, sid);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace

extern "C" {

void mi_flow_release(mi_ctx *ctx)
{
    if (!ctx || !ctx->flow) return;
    FlowState &F = *ctx->flow;
    (void)hipFree(F.params); (void)hipFree(F.m); (void)hipFree(F.v); (void)hipFree(F.grad);
    (void)hipFree(F.fwd); (void)hipFree(F.bwd);
    (void)hipFree(F.xs); (void)hipFree(F.ld); (void)hipFree(F.w); (void)hipFree(F.abar);
    (void)hipFree(F.gp); (void)hipFree(F.hist); (void)hipFree(F.acc);
    delete ctx->flow;
    ctx->flow = nullptr;
}

int mi_flow_set_params(mi_ctx *ctx, const float *params_host)
{
    if (!ctx) return MI_ERR_ARG;
    if (!params_host) return ferr(ctx, MI_ERR_ARG, "params is NULL");
    FCHK(hipSetDevice(ctx->device));
    if (!ctx->flow) ctx->flow = new FlowState();
    FlowState &F = *ctx->flow;
    if (!F.params) {
        FCHK(hipMalloc(&F.params, sizeof(float) * MI_FLOW_NPARAMS));
        FCHK(hipMalloc(&F.m, sizeof(float) * MI_FLOW_NPARAMS));
        FCHK(hipMalloc(&F.v, sizeof(float) * MI_FLOW_NPARAMS));
        FCHK(hipMalloc(&F.grad, sizeof(float) * MI_FLOW_NPARAMS));
        FCHK(hipMalloc(&F.fwd, sizeof(FwdLayer) * FL));
        FCHK(hipMalloc(&F.bwd, sizeof(BwdExtra) * FL));
        FCHK(hipMalloc(&F.acc, sizeof(double) * kRec));
    }
    FCHK(hipMemcpyAsync(F.params, params_host, sizeof(float) * MI_FLOW_NPARAMS, hipMemcpyHostToDevice, ctx->stream));
    FCHK(hipMemsetAsync(F.m, 0, sizeof(float) * MI_FLOW_NPARAMS, ctx->stream));
    FCHK(hipMemsetAsync(F.v, 0, sizeof(float) * MI_FLOW_NPARAMS, ctx->stream));
    hipLaunchKernelGGL(k_flow_pack, dim3(FL, kPackSplit), dim3(256), 0, ctx->stream, F.params, F.fwd, F.bwd);
    FCHK(hipGetLastError());
    FCHK(hipStreamSynchronize(ctx->stream));
    F.step = 0;
    F.ready = true;
    return MI_OK;
}

int mi_flow_get_params(mi_ctx *ctx, float *params_host)
{
    int rc = check_flow(ctx);
    if (rc) return rc;
    if (!params_host) return ferr(ctx, MI_ERR_ARG, "params is NULL");
    FCHK(hipSetDevice(ctx->device));
    FCHK(hipMemcpyAsync(params_host, ctx->flow->params, sizeof(float) * MI_FLOW_NPARAMS, hipMemcpyDeviceToHost,
                        ctx->stream));
    FCHK(hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

int mi_flow_init(mi_ctx *ctx, uint64_t seed)
{
    if (!ctx) return MI_ERR_ARG;
    std::vector<float> p(MI_FLOW_NPARAMS, 0.0f);
    const float dinit = (float)std::log(std::expm1(1.0 - (double)kMinD));   // softplus^-1(1 - min): identity
    uint64_t ctr = 0;
    auto uni = [&](void) {
        uint32_t o[4];
        philox_host((uint32_t)ctr, (uint32_t)(ctr >> 32), 7u, 0u, (uint32_t)seed, (uint32_t)(seed >> 32), o);
        ++ctr;
        return (float)(o[0] >> 8) * (1.0f / 16777216.0f);
    };
    for (int l = 0; l < FL; ++l) {
        float *q = p.data() + (size_t)l * FPL;
        for (int i = 0; i < 2 * FH; ++i) q[PW1 + i] = (2.0f * uni() - 1.0f) / std::sqrt(2.0f);
        for (int i = 0; i < FH * FH; ++i) q[PW2 + i] = (2.0f * uni() - 1.0f) / std::sqrt((float)FH);
        for (int i = 0; i <= FK; ++i) q[PB3 + 2 * FK + i] = dinit;
    }
    int rc = mi_flow_set_params(ctx, p.data());
    if (rc) return rc;
    if (ctx->grid.ndim != 3) {
        rc = mi_vegas_init(ctx, 3, 64);
        if (rc) return rc;
    }
    return MI_OK;
}

int mi_flow_forward(mi_ctx *ctx, uint64_t n, const float *y_dev, float *z_dev, float *logdet_dev)
{
    int rc = check_flow(ctx);
    if (rc) return rc;
    if (n == 0) return MI_OK;
    if (!y_dev || !z_dev || !logdet_dev) return ferr(ctx, MI_ERR_ARG, "NULL array");
    FCHK(hipSetDevice(ctx->device));
    FCHK(run_chain(ctx, false, false, n, y_dev, z_dev, logdet_dev, nullptr));
    FCHK(hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

int mi_flow_inverse(mi_ctx *ctx, uint64_t n, const float *z_dev, float *y_dev, float *logdet_dev)
{
    int rc = check_flow(ctx);
    if (rc) return rc;
    if (n == 0) return MI_OK;
    if (!y_dev || !z_dev || !logdet_dev) return ferr(ctx, MI_ERR_ARG, "NULL array");
    FCHK(hipSetDevice(ctx->device));
    FCHK(run_chain(ctx, true, false, n, z_dev, y_dev, logdet_dev, nullptr));
    FCHK(hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

// Backward through all layers from the layer inputs in F.xs (dL/dz = 0), then the parameter
// gradient of every layer into F.grad.  g_dev: dL/d(log q) per point, or null for the training
// step's variance-loss coefficient from F.w and the nparts per-block sums of w, w^2 in
// ctx->partials (their totals go to F.hist[step]);
// adam: the optimizer step on top.
static int launch_backward(mi_ctx *ctx, uint64_t n, const float *g_dev, uint32_t nparts, int step,
                           const AdamStep *adam)
{
    FlowState &F = *ctx->flow;
    const uint32_t blocks = bwd_blocks(ctx, n);
    const uint32_t rows = blocks;
    if (rows > F.gp_rows) return ferr(ctx, MI_ERR_STATE, "gradient partials buffer too small");
    FCHK(launch_flow_bwd(ctx->stream, blocks, F.fwd, F.bwd, n, F.xs, 3 * F.cap, g_dev, g_dev ? nullptr : F.w,
                         ctx->partials, nparts, F.abar, F.gp, F.gp_rows, g_dev ? nullptr : F.hist, step));
    const AdamStep a = adam ? *adam : AdamStep{};
    hipLaunchKernelGGL(k_grad_sum, dim3((FL * FPL + 63) / 64), dim3(256), 0, ctx->stream, F.gp, rows, F.gp_rows,
                       F.grad, a);
    FCHK(hipGetLastError());
    return MI_OK;
}

// Gradient of sum_i g_i logdet_fwd(y_i) w.r.t. the parameters, for y already in F.xs[0].
static int flow_backward_internal(mi_ctx *ctx, uint64_t n, const float *g_dev)
{
    FlowState &F = *ctx->flow;
    // forward pass keeping every layer input
    FCHK(run_chain(ctx, false, false, n, F.xs, nullptr, F.ld, F.xs));
    return launch_backward(ctx, n, g_dev, 0, 0, nullptr);
}

int mi_flow_backward(mi_ctx *ctx, uint64_t n, const float *y_dev, const float *g_dev, float *grad_dev)
{
    int rc = check_flow(ctx);
    if (rc) return rc;
    if (n == 0) return ferr(ctx, MI_ERR_ARG, "n must be > 0");
    if (!y_dev || !g_dev || !grad_dev) return ferr(ctx, MI_ERR_ARG, "NULL array");
    FCHK(hipSetDevice(ctx->device));
    rc = flow_alloc(ctx, n);
    if (rc) return rc;
    FlowState &F = *ctx->flow;
    FCHK(hipMemcpyAsync(F.xs, y_dev, sizeof(float) * 3 * n, hipMemcpyDeviceToDevice, ctx->stream));
    rc = flow_backward_internal(ctx, n, g_dev);
    if (rc) return rc;
    FCHK(hipMemcpyAsync(grad_dev, F.grad, sizeof(float) * MI_FLOW_NPARAMS, hipMemcpyDeviceToDevice, ctx->stream));
    FCHK(hipStreamSynchronize(ctx->stream));
    return MI_OK;
}

// Blocks per CU of the weight kernel's grid (the toy-PDF kernels hold 60-96 VGPRs: 5-8 waves per SIMD).
constexpr uint32_t kWeightBlocksPerCU = 6;

// The weight kernel of a run's options.
using FlowWeightKernel = void (*)(HadConst<float>, int, const float *, int, uint64_t, const float *, const float *,
                                  float *, double *, GridPdf<float>);
static FlowWeightKernel pick_flow_weight(bool grid, bool heavy, bool hw, bool stats)
{
    if (grid) {
        if (heavy) return hw ? (stats ? k_flow_weight<true, true, true, true> : k_flow_weight<true, true, false, true>)
                             : (stats ? k_flow_weight<true, false, true, true> : k_flow_weight<true, false, false, true>);
        return hw ? (stats ? k_flow_weight<false, true, true, true> : k_flow_weight<false, true, false, true>)
                  : (stats ? k_flow_weight<false, false, true, true> : k_flow_weight<false, false, false, true>);
    }
    if (heavy) return hw ? (stats ? k_flow_weight<true, true, true, false> : k_flow_weight<true, true, false, false>)
                         : (stats ? k_flow_weight<true, false, true, false> : k_flow_weight<true, false, false, false>);
    return hw ? (stats ? k_flow_weight<false, true, true, false> : k_flow_weight<false, true, false, false>)
              : (stats ? k_flow_weight<false, false, true, false> : k_flow_weight<false, false, false, false>);
}

// Sample n points of the current flow (+ VEGAS) for events ctr0 ..: y in F.xs[0], ld, w, and
// the estimate record of the batch in out_rec (device), or with out_rec null only the per-block
// sums of w and w^2 in ctx->partials (*nblocks of them, record layout).
static int flow_sample_weights(mi_ctx *ctx, const mi_run *run, uint64_t ctr0, uint64_t n, uint32_t sid,
                               double *out_rec, bool keep_all, uint32_t *nblocks = nullptr)
{
    FlowState &F = *ctx->flow;
    // z from Philox, y = T^-1(z) into F.xs (x^(0)); with keep_all every layer input as well
    FCHK(run_chain(ctx, true, true, n, nullptr, keep_all ? nullptr : F.xs, F.ld, keep_all ? F.xs : nullptr, run->key,
                   ctr0, sid));
    HadConst<float> k;
    fill_had_consts_f32(run->phys, k);
    GridPdf<float> gpdf;
    int rc = pdf_views(ctx, run->phys, &gpdf, nullptr);
    if (rc) return rc;
    const uint32_t nblk = (uint32_t)std::min<uint64_t>((n + kBlock - 1) / kBlock,
                                                      (uint64_t)(ctx->num_cu > 0 ? ctx->num_cu : 256) * kWeightBlocksPerCU);
    if ((size_t)nblk * kRec > ctx->partials_cap) return ferr(ctx, MI_ERR_STATE, "partials buffer too small");
    const size_t sm = sizeof(float) * 9 * ctx->grid.nbins;
    const bool heavy = has_heavy(run->phys.proc_mask), hw = run->precision == MI_PREC_FP32_HW, stats = !out_rec;
    const FlowWeightKernel kern = pick_flow_weight(gpdf.active != 0, heavy, hw, stats);
    hipLaunchKernelGGL(kern, dim3(nblk), dim3(kBlock), sm, ctx->stream, k, run->phys.proc_mask, ctx->grid.tab32,
                       ctx->grid.nbins, n, F.xs, F.ld, F.w, ctx->partials, gpdf);
    FCHK(hipGetLastError());
    if (nblocks) *nblocks = nblk;
    if (out_rec) FCHK(launch_reduce_records(ctx->stream, ctx->partials, nblk, out_rec));
    return MI_OK;
}

int mi_flow_train(mi_ctx *ctx, const mi_run *run, int steps, uint64_t batch, double lr, mi_flow_info *out)
{
    int rc = check_flow(ctx);
    if (rc) return rc;
    if (!run) return ferr(ctx, MI_ERR_ARG, "run is NULL");
    if (run->phys.mode != MI_MODE_HADRONIC_3D) return ferr(ctx, MI_ERR_ARG, "the flow needs the hadronic mode");
    if (steps < 1 || steps > 100000) return ferr(ctx, MI_ERR_ARG, "steps must be 1..100000");
    if (batch < 64 || batch > (1ULL << 24)) return ferr(ctx, MI_ERR_RANGE, "batch must be 64..2^24");
    if (!(lr > 0.0) || !(lr < 1.0)) return ferr(ctx, MI_ERR_ARG, "lr must be in (0, 1)");
    if (ctx->grid.ndim != 3) return ferr(ctx, MI_ERR_STATE, "the flow composes with a 3-D VEGAS grid");
    FCHK(hipSetDevice(ctx->device));
    rc = flow_alloc(ctx, batch);
    if (rc) return rc;
    FlowState &F = *ctx->flow;
    if (steps > F.hist_cap) {
        (void)hipFree(F.hist);
        F.hist = nullptr;
        FCHK(hipMalloc(&F.hist, sizeof(double) * 2 * steps));
        F.hist_cap = steps;
    }
    const uint32_t need = (uint32_t)(ctx->num_cu > 0 ? ctx->num_cu : 256) * kWeightBlocksPerCU;
    if ((size_t)need * kRec > ctx->partials_cap) {
        (void)hipFree(ctx->partials);
        ctx->partials = nullptr;
        ctx->partials_cap = 0;
        FCHK(hipMalloc(&ctx->partials, sizeof(double) * need * kRec));
        ctx->partials_cap = (size_t)need * kRec;
    }
    FCHK(hipEventRecord(ctx->ev0, ctx->stream));
    for (int s = 0; s < steps; ++s) {
        uint32_t nparts = 0;
        rc = flow_sample_weights(ctx, run, run->counter0 + (uint64_t)s * batch, batch, 2u, nullptr, true, &nparts);
        if (rc) return rc;
        // backward from the stored layer inputs (x^(0) = y ... x^(FL) = z from the sampling pass),
        // gradient and Adam's step
        F.step += 1;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.0f - std::pow(b1, (float)F.step), c2 = 1.0f - std::pow(b2, (float)F.step);
        const AdamStep adam{F.params, F.m, F.v, (float)lr, b1, b2, 1e-8f, c1, c2, F.fwd, F.bwd};
        rc = launch_backward(ctx, batch, nullptr, nparts, s, &adam);
        if (rc) return rc;
    }
    FCHK(hipEventRecord(ctx->ev1, ctx->stream));
    std::vector<double> h(2 * (size_t)steps);
    FCHK(hipMemcpyAsync(h.data(), F.hist, sizeof(double) * 2 * steps, hipMemcpyDeviceToHost, ctx->stream));
    FCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    FCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (out) {
        std::memset(out, 0, sizeof *out);
        const double nb = (double)batch;
        auto loss = [&](int s) {
            const double m = h[2 * s] / nb, m2 = h[2 * s + 1] / nb;
            return m > 0.0 ? m2 / (m * m) - 1.0 : 0.0;
        };
        out->steps = steps;
        out->loss_first = loss(0);
        out->loss_last = loss(steps - 1);
        out->cv_vegas = std::sqrt(std::max(0.0, out->loss_first));
        out->cv_flow = std::sqrt(std::max(0.0, out->loss_last));
        out->kernel_ms = ms;
    }
    for (int s = 0; s < steps; ++s)
        if (!std::isfinite(h[2 * s]) || !std::isfinite(h[2 * s + 1]))
            return ferr(ctx, MI_ERR_NUMERIC, "non-finite weights in a training batch (the flow parameters are not meaningful)");
    return MI_OK;
}

// Integration with the flow sampler: chunks of up to 2^22 events (counters counter0 ..),
// z from Philox stream 0, y = T^-1(z), weight through the VEGAS map; records accumulated on
// the device in a fixed order.
int mi_flow_integrate_internal(mi_ctx *ctx, const mi_run *run, mi_sigma_est *out)
{
    int rc = check_flow(ctx);
    if (rc) return rc;
    if (run->phys.mode != MI_MODE_HADRONIC_3D) return ferr(ctx, MI_ERR_ARG, "the flow sampler needs the hadronic mode");
    if (run->precision != MI_PREC_FP32 && run->precision != MI_PREC_FP32_HW)
        return ferr(ctx, MI_ERR_ARG, "the flow sampler evaluates FP32 weights");
    if (ctx->grid.ndim != 3) return ferr(ctx, MI_ERR_STATE, "the flow composes with a 3-D VEGAS grid");
    // chunks of 32 tiles per wave of the forward kernel's full grid (6.3M points on 256 CUs): every
    // wave walks the same number of tiles
    const uint64_t chunk = std::min<uint64_t>(run->n_points, (uint64_t)num_cu(ctx) * FWV * 64 * 32);
    rc = flow_alloc(ctx, chunk);
    if (rc) return rc;
    FlowState &F = *ctx->flow;
    const uint32_t need = (uint32_t)(ctx->num_cu > 0 ? ctx->num_cu : 256) * kWeightBlocksPerCU;
    if ((size_t)need * kRec > ctx->partials_cap) {
        (void)hipFree(ctx->partials);
        ctx->partials = nullptr;
        ctx->partials_cap = 0;
        FCHK(hipMalloc(&ctx->partials, sizeof(double) * need * kRec));
        ctx->partials_cap = (size_t)need * kRec;
    }
    FCHK(hipEventRecord(ctx->ev0, ctx->stream));
    int first = 1;
    for (uint64_t off = 0; off < run->n_points; off += chunk) {
        const uint64_t n = std::min<uint64_t>(chunk, run->n_points - off);
        rc = flow_sample_weights(ctx, run, run->counter0 + off, n, 0u, ctx->red, false);
        if (rc) return rc;
        hipLaunchKernelGGL(k_add_record, dim3(1), dim3(64), 0, ctx->stream, F.acc, ctx->red, first);
        FCHK(hipGetLastError());
        first = 0;
    }
    FCHK(hipEventRecord(ctx->ev1, ctx->stream));
    double r[kRec];
    FCHK(hipMemcpyAsync(r, F.acc, sizeof r, hipMemcpyDeviceToHost, ctx->stream));
    FCHK(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    FCHK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    std::memset(out, 0, sizeof *out);
    const double nn = (double)run->n_points;
    out->code[0] = 0;
    for (int i = 0; i < MI_MAX_PROC; ++i) out->code[1 + i] = proc_code(i);
    for (int i = 0; i <= MI_MAX_PROC; ++i) {
        if (i > 0 && !((run->phys.proc_mask >> (i - 1)) & 1)) continue;
        const double avg = r[kRS + i] / nn;
        out->sumW[i] = r[kRS + i];
        out->sumW2[i] = r[kRQ + i];
        out->wMax[i] = r[kRM + i];
        out->sigmaGen[i] = avg;
        double d2 = 0.0;
        if (avg != 0.0) d2 = (r[kRQ + i] / nn - avg * avg) / nn / (avg * avg);
        out->sigmaErr[i] = (d2 > 0.0 ? std::sqrt(d2) : 0.0) * avg;
    }
    out->nTry = out->nSel = out->nAcc = (int64_t)run->n_points;
    out->nNonZero = (int64_t)r[kRNz];
    out->kernel_ms = ms;
    if (!std::isfinite(out->sigmaGen[0])) return ferr(ctx, MI_ERR_NUMERIC, "non-finite sigma");
    return MI_OK;
}

}  // extern "C"
