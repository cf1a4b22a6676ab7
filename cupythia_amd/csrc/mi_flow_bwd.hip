// mi_flow_bwd.hip -- backward kernel of the spline-coupling flow: the adjoint of one layer for
// L = sum_i (abar_out_i . x_out_i + g_i log S'_i) summed with the later layers' contributions.
// Per layer, abar (n x 3) holds dL/d(output coordinates) on entry and dL/d(input coordinates) on
// exit; the layer's parameter gradient of the block's points goes to one row of gp.
//
// A wave owns 64 points per tile and keeps every matrix of the tile in matrix-core registers
// (layout in mi_flow_common.h): the MLP is recomputed as in the forward kernel (activations H1,
// H2 with rows = units, columns = points), the spline adjoint runs per point in registers, and
// dH2 = W3^T dO, dH1 = W2^T dZ2 and dc = W1^T dZ1 are products with the transposed weights.  The
// weight gradients dW3 = dO^T H2, dW2 = dZ2^T H1 and (dW1, db1) = dZ1^T (c, 1) contract over the
// points, so their operands are transposed first -- exactly, as products with one-hot fragments
// on the matrix cores -- and accumulate across the wave's tiles in the accumulation registers
// (AGPRs; one wave per SIMD) while every other product stays in ordinary VGPRs, where the VALU
// work reads it without copies.
#include "mi_flow_common.h"

namespace mi {

namespace {

// LDS: the layer's weight images and, per wave, the conditioners' BF16 parts; after the tiles the
// block's FWB gradient rows take its place.
struct BwdTiles {
    FwdLayer f;
    BwdExtra b;
    uint16_t comp[FWB][6][64];   // [wave][part][point]: c0 hi, mid, lo, c1 hi, mid, lo
};
union BwdShared {
    BwdTiles t;
    float rows[FWB][FPLP];
};
static_assert(sizeof(BwdShared) <= 160 * 1024, "backward LDS");
static_assert(offsetof(BwdTiles, b) == sizeof(FwdLayer), "the two weight images stage as one array");
constexpr int kStageF = sizeof(FwdLayer) / 16, kStageB = sizeof(BwdExtra) / 16;
constexpr int kStageIt = (kStageF + kStageB + 64 * FWB - 1) / (64 * FWB);   // 16-byte loads per work-item

// 0xFFFF where the BF16 activation (ReLU output: 0 .. 0x7FFF as a 16-bit integer) is zero, per
// half: h - 1 is negative exactly there.
__device__ __forceinline__ uint32_t zero_mask(uint32_t a)
{
    const s16x2 v = __builtin_bit_cast(s16x2, a) - (s16x2){1, 1};
    return __builtin_bit_cast(uint32_t, v >> (s16x2){15, 15});
}
// Operand of K step kk of a block, zero where the activation fragment h is zero (ReLU derivative).
__device__ __forceinline__ bf16x8 masked_frag(const f32x16 &c, int base, bf16x8 h)
{
    const uint4 m = __builtin_bit_cast(uint4, h);
    uint4 w;
    w.x = pack_bf16x2(c[base], c[base + 1]) & ~zero_mask(m.x);
    w.y = pack_bf16x2(c[base + 2], c[base + 3]) & ~zero_mask(m.y);
    w.z = pack_bf16x2(c[base + 4], c[base + 5]) & ~zero_mask(m.z);
    w.w = pack_bf16x2(c[base + 6], c[base + 7]) & ~zero_mask(m.w);
    return __builtin_bit_cast(bf16x8, w);
}

// Exact transpose of an operand set X[kk][pb] (as A operands: rows = points of block pb, K = the
// source rows in swz order): the products with one-hot fragments give blocks T[pb][ub] with points
// in registers and source rows 32 ub + l % 32 on lanes; Xt[kk][ub] are their operands with the
// points as the contraction index (slot k carries point swz(k)).
__device__ __forceinline__ void transpose64(const bf16x8 (&X)[4][2], const bf16x8 (&I)[4], f32x16 (&T)[2][2],
                                            bf16x8 (&Xt)[4][2])
{
    const f32x16 zero = {};
#pragma unroll
    for (int ub = 0; ub < 2; ++ub) {
#pragma unroll
        for (int pb = 0; pb < 2; ++pb) {
            T[pb][ub] = mfma16(X[2 * ub][pb], I[2 * ub], zero);
            T[pb][ub] = mfma16(X[2 * ub + 1][pb], I[2 * ub + 1], T[pb][ub]);
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) Xt[kk][ub] = plain_frag(T[kk >> 1][ub], 8 * (kk & 1));
    }
}
// acc += a b with acc in the accumulation registers, where it stays across the tiles (the other
// products run in the VGPR form, compiler flag -amdgpu-mfma-vgpr-form).  a and b are always
// written by VALU instructions or LDS loads; the five wait states before the product cover a
// VALU write just before (two are required).
typedef int i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void mfma_acc(f32x16 &acc, bf16x8 a, bf16x8 b)
{
    asm("s_nop 4\n\tv_mfma_f32_32x32x16_bf16 %0, %1, %2, %0"
        : "+a"(acc)
        : "v"(__builtin_bit_cast(i32x4, a)), "v"(__builtin_bit_cast(i32x4, b)));
}
// Waits until the last products into acc have been written (more than the 18 wait states a
// 16-pass product needs before its result is read).
__device__ __forceinline__ void mfma_acc_drain(f32x16 &acc)
{
    asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7" : "+a"(acc));
}

// Sum over the points a lane holds in registers (half of the tile's points; the lane l ^ 32 holds
// the other half).
__device__ __forceinline__ float reg_sum(const f32x16 &a, const f32x16 &b)
{
    float s = 0.0f;
#pragma unroll
    for (int v = 0; v < 16; ++v) s += a[v] + b[v];
    return s;
}

// Adjoint of the forward spline of one point: given its MLP outputs o, the input coordinate x,
// aS = dL/dy and gL = dL/d(log S'), returns dL/dx and writes dO = dL/do.  The forward map is the
// one of the forward kernel: widths and heights kMin + (1 - FK kMin) softmax(o), knot derivatives
// kMin + softplus(o), the bin of x located on the cumulative widths.
__device__ __forceinline__ float spline_bwd(const float (&o)[FO], float x, float aS, float gL, float (&dO)[FO])
{
    constexpr float L2E = 1.4426950408889634f;
    const float cw = 1.0f - FK * kMinW, ch = 1.0f - FK * kMinH;
    float mw = o[0], mh = o[FK];
#pragma unroll
    for (int j = 1; j < FK; ++j) { mw = fmaxf(mw, o[j]); mh = fmaxf(mh, o[FK + j]); }
    float pw[FK], ph[FK];
    const float nw = -mw * L2E, nh = -mh * L2E;
    float Sw = 0.0f, Sh = 0.0f;
#pragma unroll
    for (int j = 0; j < FK; ++j) {
        pw[j] = exp2_hw(__builtin_fmaf(o[j], L2E, nw));
        ph[j] = exp2_hw(__builtin_fmaf(o[FK + j], L2E, nh));
        Sw += pw[j];
        Sh += ph[j];
    }
    const float iSw = rcp_hw(Sw), iSh = rcp_hw(Sh);
    const float rw = cw * iSw;
    // bin k of x on the width knots kMinW j + cw Cw_j / Sw (Cw_j, Ch_j: the running sums below j),
    // with the quantities of bin k
    float sX = 0.0f, sCw = 0.0f, sew = pw[0], sCh = 0.0f, seh = ph[0], sk = 0.0f, sd0 = o[2 * FK], sd1 = o[2 * FK + 1];
    float Cw = 0.0f, Ch = 0.0f;
#pragma unroll
    for (int j = 1; j < FK; ++j) {
        Cw += pw[j - 1];
        Ch += ph[j - 1];
        const float K = __builtin_fmaf(rw, Cw, kMinW * (float)j);
        const bool c = x >= K;
        sX = c ? K : sX;
        sCw = c ? Cw : sCw;
        sew = c ? pw[j] : sew;
        sCh = c ? Ch : sCh;
        seh = c ? ph[j] : seh;
        sk = c ? (float)j : sk;
        sd0 = c ? o[2 * FK + j] : sd0;
        sd1 = c ? o[2 * FK + j + 1] : sd1;
    }
#pragma unroll
    for (int j = 0; j < FK; ++j) { pw[j] *= iSw; ph[j] *= iSh; }
    const float pwk = sew * iSw, phk = seh * iSh;                 // softmax weights of bin k
    const float xk = sX, wk = __builtin_fmaf(cw, pwk, kMinW);
    const float hk = __builtin_fmaf(ch, phk, kMinH);
    float sg0, sg1;
    const float dk = kMinD + softplus_hw(sd0, &sg0), dk1 = kMinD + softplus_hw(sd1, &sg1);
    // the map y = yk + hk R(xi, s, dk, dk1) and log S' = log(s^2 n2 / den^2)
    const float iw = rcp_hw(wk);
    const float xi = (x - xk) * iw, s = hk * iw, t = xi * (1.0f - xi);
    const float A = dk1 + dk - 2.0f * s, den = __builtin_fmaf(A, t, s);
    const float iden = rcp_hw(den), iden2 = iden * iden;
    const float P2 = __builtin_fmaf(s * xi, xi, dk * t), R = P2 * iden;
    const float omx = 1.0f - xi, om2t = 1.0f - 2.0f * t, om2x = 1.0f - 2.0f * xi;
    const float n2 = __builtin_fmaf(dk1 * xi, xi, __builtin_fmaf(2.0f * s, t, dk * omx * omx));
    const float in2 = rcp_hw(n2);
    // d R / d (xi, s, dk, dk1)
    const float Pxi = __builtin_fmaf(2.0f * s, xi, dk * om2x), denxi = A * om2x;
    const float R_xi = (Pxi * den - P2 * denxi) * iden2;
    const float R_s = (xi * xi * den - P2 * om2t) * iden2;
    const float R_dk = t * (den - P2) * iden2;
    const float R_dk1 = -P2 * t * iden2;
    // d log S' / d (xi, s, dk, dk1)
    const float n2xi = 2.0f * dk1 * xi + 2.0f * s * om2x - 2.0f * dk * omx;
    const float L_xi = n2xi * in2 - 2.0f * denxi * iden;
    const float L_s = 2.0f * rcp_hw(s) + 2.0f * t * in2 - 2.0f * om2t * iden;
    const float L_dk = omx * omx * in2 - 2.0f * t * iden;
    const float L_dk1 = xi * xi * in2 - 2.0f * t * iden;
    const float hR = aS * hk;
    const float Gxi = hR * R_xi + gL * L_xi, Gs = hR * R_s + gL * L_s;
    const float Gdk = hR * R_dk + gL * L_dk, Gdk1 = hR * R_dk1 + gL * L_dk1;
    const float dx = Gxi * iw;                                      // dL / dx
    const float dxk = -dx, dwk = -(Gxi * xi + Gs * s) * iw, dhk = aS * R + Gs * iw, dyk = aS;
    // widths / heights: bin k's start is the sum of the softmax weights below k, its size the
    // weight of k; then the softmax Jacobian p_q (g_q - sum_r p_r g_r)
    const float dotw = cw * (dxk * (sCw * iSw) + dwk * pwk), doth = ch * (dyk * (sCh * iSh) + dhk * phk);
    const float aw_lt = cw * dxk - dotw, aw_eq = cw * dwk - dotw, aw_gt = -dotw;
    const float ah_lt = ch * dyk - doth, ah_eq = ch * dhk - doth, ah_gt = -doth;
#pragma unroll
    for (int q = 0; q < FK; ++q) {
        const bool lt = (float)q < sk, eq = (float)q == sk;
        dO[q] = pw[q] * (lt ? aw_lt : (eq ? aw_eq : aw_gt));
        dO[FK + q] = ph[q] * (lt ? ah_lt : (eq ? ah_eq : ah_gt));
    }
    const float gd0 = Gdk * sg0, gd1 = Gdk1 * sg1;
#pragma unroll
    for (int q = 0; q <= FK; ++q) dO[2 * FK + q] = ((float)q == sk ? gd0 : 0.0f) + ((float)q == sk + 1.0f ? gd1 : 0.0f);
    return dx;
}

// Arguments of the backward kernel.
struct BwdArgs {
    const FwdLayer *fl;
    const BwdExtra *bl;
    uint64_t n;
    const float *xs;            // layer inputs of the normalizing direction: x^(l) at xs + l * xstride
    uint64_t xstride;
    const float *g;             // dL/d(log q) per point, or null: the variance-loss coefficient
    const float *w;             //   g_i = -(w_i / <w>)^2 / n, <w> = sum w / n from the nparts
    const double *parts;        //   per-block sums of w and w^2 (record slots kRS, kRQ)
    uint32_t nparts;
    float *abar;                // n x 3 work array (not read at the last layer, where dL/dz = 0)
    float *gp;                  // gradient rows of layer l: gp + (l * gp_stride + blockIdx.x) * FPLP
    uint32_t gp_stride;
    double *hist;               // optional: (sum w, sum w^2) of the batch at hist[2 step]
    int step;
};

// All layers, last to first, in one launch: per layer the block stages the layer's weights, its
// waves walk their tiles (the same tiles in every layer, so that abar passes from layer to layer
// within a thread), and the block writes its gradient row.
__global__ __launch_bounds__(64 * FWB) void k_flow_bwd(const BwdArgs A)
{
    __shared__ BwdShared S;
    const FwdLayer &W = S.t.f;
    const BwdExtra &T = S.t.b;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, hh = lane >> 5, c = lane & 31;
    uint16_t(*comp)[64] = S.t.comp[wave];
    // the batch's sum w and sum w^2, the same in every block (fixed order: strided per thread, the
    // wave's tree, the waves in order)
    double inv_mean = 1.0;
    if (!A.g) {
        __shared__ double sums[2][FWB];
        double sw = 0.0, sq = 0.0;
        for (uint32_t b = threadIdx.x; b < A.nparts; b += blockDim.x) {
            sw += A.parts[(size_t)b * kRec + kRS];
            sq += A.parts[(size_t)b * kRec + kRQ];
        }
        sw = wave_sum(sw);
        sq = wave_sum(sq);
        if (lane == 0) { sums[0][wave] = sw; sums[1][wave] = sq; }
        __syncthreads();
        sw = sums[0][0];
        sq = sums[1][0];
#pragma unroll
        for (int w = 1; w < FWB; ++w) { sw += sums[0][w]; sq += sums[1][w]; }
        if (A.hist && blockIdx.x == 0 && threadIdx.x == 0) {
            A.hist[2 * A.step] = sw;
            A.hist[2 * A.step + 1] = sq;
        }
        inv_mean = (double)A.n / sw;
    }
    // g_i = -(w_i / <w>)^2 / n as -(w_i * inv_mean)^2 * inv_n
    const double inv_n = 1.0 / (double)A.n;
    // one-hot fragments of the transposes: step kk, column 32 (kk / 2) + c, a one in the slot
    // that carries row 32 (kk / 2) + c
    bf16x8 I[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (swz(16 * kk + 8 * hh + j) == 32 * (kk >> 1) + c) w[j >> 1] |= 0x3F80u << (16 * (j & 1));
        I[kk] = as_frag(make_uint4(w[0], w[1], w[2], w[3]));
    }
    const f32x16 zero = {};
    const uint64_t ntiles = (A.n + 63) / 64;
#pragma unroll 1
    for (int l = FL - 1; l >= 0; --l) {
        {
            // The layer's two weight images (contiguous in LDS), every 16-byte load of the work-item
            // in flight before its LDS writes: one wave per SIMD has no other wave to hide a
            // load-then-write loop's latency.  The index is an opaque copy, so that the
            // addresses are formed here and not held across the tiles.
            int tid = threadIdx.x;
            __asm__ volatile("" : "+v"(tid));
            const uint4 *s0 = reinterpret_cast<const uint4 *>(A.fl + l);
            const uint4 *s1 = reinterpret_cast<const uint4 *>(A.bl + l);
            uint4 *d = reinterpret_cast<uint4 *>(&S.t.f);
            const auto ld = [=](int k) {
                const int j = tid + 64 * FWB * k;
                uint4 r = make_uint4(0u, 0u, 0u, 0u);
                if (j < kStageF + kStageB) r = j < kStageF ? s0[j] : s1[j - kStageF];
                return r;
            };
            const auto st = [=](int k, uint4 r) {
                const int j = tid + 64 * FWB * k;
                if (j < kStageF + kStageB) d[j] = r;
            };
            // named values, not an array: a private array is placed in LDS before it is unrolled
            static_assert(kStageIt == 9, "nine 16-byte loads per work-item");
            const uint4 v0 = ld(0), v1 = ld(1), v2 = ld(2), v3 = ld(3), v4 = ld(4), v5 = ld(5), v6 = ld(6),
                        v7 = ld(7), v8 = ld(8);
            st(0, v0); st(1, v1); st(2, v2); st(3, v3); st(4, v4); st(5, v5); st(6, v6); st(7, v7); st(8, v8);
            __syncthreads();
        }
        const int d = l % 3, ca = cond0(d), cb = cond1(d);
        const float *xin = A.xs + (size_t)l * A.xstride;
        float *abar = A.abar;
        f32x16 gW3[2][2], gW2[2][2], gW1[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            gW1[a] = zero;
#pragma unroll
            for (int b = 0; b < 2; ++b) { gW3[a][b] = zero; gW2[a][b] = zero; }
        }
        float gb2[2] = {0.0f, 0.0f}, gb3[2] = {0.0f, 0.0f};
        for (uint64_t tile = (uint64_t)blockIdx.x * FWB + wave; tile < ntiles; tile += (uint64_t)gridDim.x * FWB) {
            const uint64_t i = tile * 64 + lane;
            const bool live = i < A.n;
            float x[3] = {0.5f, 0.5f, 0.5f}, ab[3] = {0.0f, 0.0f, 0.0f}, gi = 0.0f;
            if (live) {
                x[0] = xin[3 * i]; x[1] = xin[3 * i + 1]; x[2] = xin[3 * i + 2];
                if (l < FL - 1) { ab[0] = abar[3 * i]; ab[1] = abar[3 * i + 1]; ab[2] = abar[3 * i + 2]; }
                if (A.g) {
                    gi = A.g[i];
                } else {
                    const double r = (double)A.w[i] * inv_mean;
                    gi = (float)(-(r * r) * inv_n);
                }
            }
            const CondFrag cf = cond_frag(pick3(x, ca), pick3(x, cb));
            wave_sync();                                   // the previous tile's reads of comp are done
            comp[0][lane] = (uint16_t)cf.parts[0]; comp[3][lane] = (uint16_t)(cf.parts[0] >> 16);
            comp[1][lane] = (uint16_t)cf.parts[1]; comp[4][lane] = (uint16_t)(cf.parts[1] >> 16);
            comp[2][lane] = (uint16_t)cf.parts[2]; comp[5][lane] = (uint16_t)(cf.parts[2] >> 16);
            // ---- forward recompute: H1, H2 as operands (rows = units, columns = points), outputs o
            bf16x8 H1[4][2], H2[4][2];
            f32x16 C[2][2];
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                const bf16x8 a = as_frag(W.W1f[hh][32 * mb + c]);
                C[mb][0] = mfma16(a, as_frag(cf.frag[0]), zero);
                C[mb][1] = mfma16(a, as_frag(cf.frag[1]), zero);
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) H1[kk][nb] = relu_frag(C[kk >> 1][nb], 8 * (kk & 1));
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                C[mb][0] = bias_block(W.b2, mb, hh);
                C[mb][1] = C[mb][0];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const bf16x8 a = as_frag(W.W2f[kk][hh][32 * mb + c]);
                    C[mb][0] = mfma16(a, H1[kk][0], C[mb][0]);
                    C[mb][1] = mfma16(a, H1[kk][1], C[mb][1]);
                }
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) H2[kk][nb] = relu_frag(C[kk >> 1][nb], 8 * (kk & 1));
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                C[mb][0] = bias_block(W.b3, mb, hh);
                C[mb][1] = C[mb][0];
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const bf16x8 a = as_frag(W.W3f[kk][hh][32 * mb + c]);
                    C[mb][0] = mfma16(a, H2[kk][0], C[mb][0]);
                    C[mb][1] = mfma16(a, H2[kk][1], C[mb][1]);
                }
            }
            // ---- spline adjoint of the lane's point, dO back to the block layout
            float dO[FO];
            float dx;
            {
                float o[FO];
                rows_to_lane(C, o);
                dx = spline_bwd(o, pick3(x, d), live ? pick3(ab, d) : 0.0f, live ? gi : 0.0f, dO);
            }
            lane_to_rows(dO, C);
            bf16x8 Df[4][2];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) Df[kk][nb] = plain_frag(C[kk >> 1][nb], 8 * (kk & 1));
            // ---- dW3 += dO^T H2, db3 += sum dO
            {
                f32x16 Tt[2][2];
                bf16x8 Ot[4][2], Ht[4][2];
                transpose64(Df, I, Tt, Ot);
                gb3[0] += reg_sum(Tt[0][0], Tt[1][0]);
                gb3[1] += reg_sum(Tt[0][1], Tt[1][1]);
                transpose64(H2, I, Tt, Ht);
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                    for (int ob = 0; ob < 2; ++ob)
#pragma unroll
                        for (int ub = 0; ub < 2; ++ub) mfma_acc(gW3[ob][ub], Ot[kk][ob], Ht[kk][ub]);
            }
            // ---- dZ2 = (W3^T dO) where H2 > 0
            bf16x8 Z2[4][2];
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                C[mb][0] = zero;
                C[mb][1] = zero;
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const bf16x8 a = as_frag(T.W3Tf[kk][hh][32 * mb + c]);
                    C[mb][0] = mfma16(a, Df[kk][0], C[mb][0]);
                    C[mb][1] = mfma16(a, Df[kk][1], C[mb][1]);
                }
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) Z2[kk][nb] = masked_frag(C[kk >> 1][nb], 8 * (kk & 1), H2[kk][nb]);
            // ---- dW2 += dZ2^T H1, db2 += sum dZ2
            {
                f32x16 Tt[2][2];
                bf16x8 Zt[4][2], Ht[4][2];
                transpose64(Z2, I, Tt, Zt);
                gb2[0] += reg_sum(Tt[0][0], Tt[1][0]);
                gb2[1] += reg_sum(Tt[0][1], Tt[1][1]);
                transpose64(H1, I, Tt, Ht);
#pragma unroll
                for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                    for (int ub = 0; ub < 2; ++ub)
#pragma unroll
                        for (int kb = 0; kb < 2; ++kb) mfma_acc(gW2[ub][kb], Zt[kk][ub], Ht[kk][kb]);
            }
            // ---- dZ1 = (W2^T dZ2) where H1 > 0
            bf16x8 Z1[4][2];
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                C[mb][0] = zero;
                C[mb][1] = zero;
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const bf16x8 a = as_frag(T.W2Tf[kk][hh][32 * mb + c]);
                    C[mb][0] = mfma16(a, Z2[kk][0], C[mb][0]);
                    C[mb][1] = mfma16(a, Z2[kk][1], C[mb][1]);
                }
            }
#pragma unroll
            for (int kk = 0; kk < 4; ++kk)
#pragma unroll
                for (int nb = 0; nb < 2; ++nb) Z1[kk][nb] = masked_frag(C[kk >> 1][nb], 8 * (kk & 1), H1[kk][nb]);
            // ---- (dW1, db1) += dZ1^T (c0 parts, c1 parts, 1): column q < 6 of the operand is part q of
            // the points' conditioners, column 6 is one
            {
                f32x16 Tt[2][2];
                bf16x8 Zt[4][2];
                transpose64(Z1, I, Tt, Zt);
                wave_sync();                               // comp written by the whole wave
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    uint4 w = make_uint4(0u, 0u, 0u, 0u);
                    if (c < 6) {
                        const uint2 a = *reinterpret_cast<const uint2 *>(&comp[c][16 * kk + 4 * hh]);
                        const uint2 b = *reinterpret_cast<const uint2 *>(&comp[c][16 * kk + 8 + 4 * hh]);
                        w = make_uint4(a.x, a.y, b.x, b.y);
                    } else if (c == 6) {
                        w = make_uint4(0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u);
                    }
#pragma unroll
                    for (int ub = 0; ub < 2; ++ub) mfma_acc(gW1[ub], Zt[kk][ub], as_frag(w));
                }
            }
            // ---- dc = W1^T dZ1: rows 0, 1 of the product, in lanes 0..31 for the points of each block
            float dc0, dc1;
            {
                f32x16 E[2] = {zero, zero};
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    const bf16x8 a = c < 2 ? as_frag(T.W1Tf[kk][hh][c & 1]) : as_frag(make_uint4(0u, 0u, 0u, 0u));
                    E[0] = mfma16(a, Z1[kk][0], E[0]);
                    E[1] = mfma16(a, Z1[kk][1], E[1]);
                }
                dc0 = __uint_as_float(swap32(__float_as_uint(E[0][0]), __float_as_uint(E[1][0])).a);
                dc1 = __uint_as_float(swap32(__float_as_uint(E[0][1]), __float_as_uint(E[1][1])).a);
            }
            if (live) {
                set3(ab, d, dx);
                set3(ab, ca, pick3(ab, ca) + dc0);
                set3(ab, cb, pick3(ab, cb) + dc1);
                abar[3 * i] = ab[0]; abar[3 * i + 1] = ab[1]; abar[3 * i + 2] = ab[2];
            }
        }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            mfma_acc_drain(gW1[a]);
#pragma unroll
            for (int b = 0; b < 2; ++b) { mfma_acc_drain(gW3[a][b]); mfma_acc_drain(gW2[a][b]); }
        }
        // the lanes l and l ^ 32 hold the two halves of the points
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            gb2[h] += __shfl_xor(gb2[h], 32, 64);
            gb3[h] += __shfl_xor(gb3[h], 32, 64);
        }
        __syncthreads();                                   // every wave is past its tiles: LDS holds the rows now
        float *out = S.rows[wave];
        // dW1 and db1: columns 0..2 (c0 parts), 3..5 (c1 parts) and 6 (ones) of each row, summed
        // per row through the W2 part of the row before that is written
        {
            float *tmp = out + PW2;
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int v = 0; v < 16; ++v) tmp[(16 * a + v) * 64 + lane] = gW1[a][v];
            wave_sync();
            const int u = lane;                                  // row u: block u / 32, register v, half h
            const int a = u >> 5, v = 4 * ((u >> 3) & 3) + (u & 3), h = (u >> 2) & 1;
            const float *q = tmp + (16 * a + v) * 64 + 32 * h;
            const float w0 = (q[0] + q[1]) + q[2], w1 = (q[3] + q[4]) + q[5], b1 = q[6];
            wave_sync();
            out[PW1 + 2 * u] = w0;
            out[PW1 + 2 * u + 1] = w1;
            out[PB1 + u] = b1;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int row = 32 * a + 8 * (v >> 2) + 4 * hh + (v & 3), col = 32 * b + c;
                    if (row < FO) out[PW3 + row * FH + col] = gW3[a][b][v];
                    out[PW2 + row * FH + col] = gW2[a][b][v];
                }
        if (hh == 0) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                out[PB2 + 32 * h + c] = gb2[h];
                if (32 * h + c < FO) out[PB3 + 32 * h + c] = gb3[h];
            }
        }
        __syncthreads();
        // the block's row, four floats at a time (the three of the padding carry no parameter)
        float4 *dst = reinterpret_cast<float4 *>(A.gp + ((size_t)l * A.gp_stride + blockIdx.x) * FPLP);
        for (int j = threadIdx.x; j < FPLP / 4; j += 64 * FWB) {
            float4 t = reinterpret_cast<const float4 *>(S.rows[0])[j];
#pragma unroll
            for (int w = 1; w < FWB; ++w) {
                const float4 u = reinterpret_cast<const float4 *>(S.rows[w])[j];
                t.x += u.x; t.y += u.y; t.z += u.z; t.w += u.w;
            }
            dst[j] = t;
        }
        __syncthreads();                                   // rows read before the next layer's weights replace them
    }
}

}  // namespace

hipError_t launch_flow_bwd(hipStream_t stream, uint32_t blocks, const FwdLayer *fl, const BwdExtra *bl, uint64_t n,
                           const float *xs, uint64_t xstride, const float *g, const float *w, const double *parts,
                           uint32_t nparts, float *abar, float *gp, uint32_t gp_stride, double *hist, int step)
{
    const BwdArgs a{fl, bl, n, xs, xstride, g, w, parts, nparts, abar, gp, gp_stride, hist, step};
    hipLaunchKernelGGL(k_flow_bwd, dim3(blocks), dim3(64 * FWB), 0, stream, a);
    return hipGetLastError();
}

}  // namespace mi
