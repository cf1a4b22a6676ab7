// mi_flow_common.h -- definitions shared by the flow's forward / inverse kernel (mi_flow.hip) and
// its backward kernel (mi_flow_bwd.hip): parameter layout, the weight images the kernels stage in
// LDS, and the register-level building blocks of the matrix-core MLP.
//
// Matrix-core layout (v_mfma_f32_32x32x16_bf16, wave64): for C = A B with a 32 x 16 A block and a
// 16 x 32 B block, lane l holds A[l % 32][8 (l / 32) .. + 7] and B[8 (l / 32) .. + 7][l % 32];
// the result element (m, n) sits in lane n + 32 ((m / 4) % 2), register 4 (m / 8) + m % 4.  A
// 64 x 64 matrix X is kept as four such blocks (mb, nb): lane l, register v = 4 i + r of block
// (mb, nb) holds X[32 mb + 8 i + 4 (l / 32) + r][32 nb + l % 32] -- rows in registers, columns on
// lanes.  Eight consecutive registers of a block are an MFMA operand of the next product as they
// stand, once the contraction index is permuted: slot k of a 16-wide K step carries row swz(k)
// (bits 2 and 3 of k exchanged).  Products of two operands built that way contract over the
// source rows; weights are stored with the same permutation.
#pragma once

#include "mi_kernel_common.h"

namespace mi {

constexpr int FL = MI_FLOW_LAYERS, FK = MI_FLOW_BINS, FH = MI_FLOW_HIDDEN, FO = MI_FLOW_NOUT;
constexpr int FPL = MI_FLOW_NPARAMS / MI_FLOW_LAYERS;   // 7537 parameters per layer
constexpr int FPLP = (FPL + 3) & ~3;                    // row stride of the gradient rows (16-byte aligned)
constexpr int PW1 = 0, PB1 = 128, PW2 = 192, PB2 = 4288, PW3 = 4352, PB3 = 7488;
static_assert(PB3 + FO == FPL, "flow parameter layout");
constexpr float kMinW = 1e-3f, kMinH = 1e-3f, kMinD = 1e-3f;
constexpr int FWV = 12;     // waves per block of the forward / inverse kernel (one block per CU)
constexpr int FWB = 4;      // waves per block of the backward kernel (one block per CU, one wave per SIMD)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));

__host__ __device__ constexpr int swz(int k) { return (k & ~12) | ((k & 4) << 1) | ((k & 8) >> 1); }

// Weights of one layer as MFMA operand fragments, one 16-byte fragment per lane, laid out
// [K step][half][row] so that a wave's fragment load reads 1 KB of consecutive LDS.  As A
// operands (rows = output units) they give W X; the same fragments as B operands (columns =
// output units) give X^T W^T.
struct FwdLayer {
    uint4 W1f[2][64];        // [0][m]: (W1[m][0], W1[m][1]) x 3, 0, 0; [1][m]: b1[m] in 3 BF16 parts, 0 x 5
    uint4 W2f[4][2][64];     // [kk][hh][m]: W2[m][swz(16 kk + 8 hh + j)], j = 0..7
    uint4 W3f[4][2][64];     // the same for W3; rows 49..63 zero
    float b2[64], b3[64];    // b3: 49 used, rest 0
};
static_assert(sizeof(FwdLayer) % 16 == 0, "FwdLayer copies as uint4");

// The transposed weights the backward kernel needs (products W^T D).
struct BwdExtra {
    uint4 W2Tf[4][2][64];    // [kk][hh][u]: W2[swz(16 kk + 8 hh + j)][u]
    uint4 W3Tf[4][2][64];    // [kk][hh][u]: W3[swz(16 kk + 8 hh + j)][u], zero where the row is >= 49
    uint4 W1Tf[4][2][2];     // [kk][hh][c]: W1[swz(16 kk + 8 hh + j)][c]
};
static_assert(sizeof(BwdExtra) % 16 == 0, "BwdExtra copies as uint4");

// FP32 -> BF16, round to nearest even (gfx950's v_cvt_pk_bf16_f32 converts two at a time).
__device__ __forceinline__ uint32_t pack_bf16x2(float a, float b)
{
    const f32x2 v = {a, b};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}
__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }
__device__ __forceinline__ bf16x8 as_frag(uint4 w) { return __builtin_bit_cast(bf16x8, w); }

__device__ __forceinline__ f32x16 mfma16(bf16x8 a, bf16x8 b, f32x16 c)
{
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Orders the LDS accesses of one wave around a point where lanes start reading what other lanes
// of the same wave wrote (the LDS executes one wave's operations in issue order; this keeps the
// compiler from moving them across).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// bf16(relu(a)), bf16(relu(b)) packed: ReLU on the BF16 bit patterns as a signed 16-bit maximum
// with 0 (negative values and -0 have the sign bit set).
__device__ __forceinline__ uint32_t relu_bf16x2(float a, float b)
{
    const s16x2 v = __builtin_bit_cast(s16x2, pack_bf16x2(a, b));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(v, (s16x2){0, 0}));
}
// Operand of K step kk from the block holding rows 32 (kk / 2) ..: registers 8 (kk % 2) .. + 7,
// with (relu_frag) or without (plain_frag) the ReLU.
__device__ __forceinline__ bf16x8 relu_frag(const f32x16 &c, int base)
{
    uint4 w;
    w.x = relu_bf16x2(c[base], c[base + 1]);
    w.y = relu_bf16x2(c[base + 2], c[base + 3]);
    w.z = relu_bf16x2(c[base + 4], c[base + 5]);
    w.w = relu_bf16x2(c[base + 6], c[base + 7]);
    return __builtin_bit_cast(bf16x8, w);
}
__device__ __forceinline__ bf16x8 plain_frag(const f32x16 &c, int base)
{
    uint4 w;
    w.x = pack_bf16x2(c[base], c[base + 1]);
    w.y = pack_bf16x2(c[base + 2], c[base + 3]);
    w.z = pack_bf16x2(c[base + 4], c[base + 5]);
    w.w = pack_bf16x2(c[base + 6], c[base + 7]);
    return __builtin_bit_cast(bf16x8, w);
}
// Block of rows 32 mb .. + 31 started at the bias: register 4 i + r = b[32 mb + 8 i + 4 hh + r].
__device__ __forceinline__ f32x16 bias_block(const float *b, int mb, int hh)
{
    f32x16 c;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float4 v = *reinterpret_cast<const float4 *>(b + 32 * mb + 8 * i + 4 * hh);
        c[4 * i] = v.x; c[4 * i + 1] = v.y; c[4 * i + 2] = v.z; c[4 * i + 3] = v.w;
    }
    return c;
}
// v_permlane32_swap(a, b): the upper half of a's lanes trades places with the lower half of b's.
struct Swapped { uint32_t a, b; };
__device__ __forceinline__ Swapped swap32(uint32_t a, uint32_t b)
{
    const auto t = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    return {t[0], t[1]};
}
// The value of lane l ^ 32 in lanes 0..31 (lanes 32..63 keep their own).
__device__ __forceinline__ uint32_t swap_lo(uint32_t v) { return swap32(v, v).b; }

// The conditioners (c0, c1) of the point of this lane as the K = 16 operand of the first layer:
// lanes 0..31 (K slots 0..7) the three BF16 parts (hi, mid, lo; exact) of c0 and c1 interleaved,
// lanes 32..63 (slots 8..15) ones against b1's three parts.  frag[0] covers points 0..31, frag[1]
// points 32..63; parts[] are the packed (c0, c1) pairs hi, mid, lo of the lane's own point.
struct CondFrag { uint4 frag[2]; uint32_t parts[3]; };
__device__ __forceinline__ CondFrag cond_frag(float c0, float c1)
{
    const int hh = (threadIdx.x & 63) >> 5;
    CondFrag f;
    const uint32_t p = pack_bf16x2(c0, c1);
    const float e0 = c0 - __uint_as_float(p << 16), e1 = c1 - __uint_as_float(p & 0xFFFF0000u);
    const uint32_t q = pack_bf16x2(e0, e1);
    const uint32_t s = pack_bf16x2(e0 - __uint_as_float(q << 16), e1 - __uint_as_float(q & 0xFFFF0000u));
    const uint32_t ps = swap_lo(p), qs = swap_lo(q), ss = swap_lo(s);
    const uint4 ones = make_uint4(0x3F803F80u, 0x3F80u, 0u, 0u);
    f.frag[0] = hh == 0 ? make_uint4(p, q, s, 0u) : ones;
    f.frag[1] = hh == 0 ? make_uint4(ps, qs, ss, 0u) : ones;
    f.parts[0] = p; f.parts[1] = q; f.parts[2] = s;
    return f;
}

// Block pair (mb, nb = 0, 1) of a matrix with rows = MLP outputs and columns = points -> the FO
// outputs of the lane's own point: register 4 i + r of (mb, 0) and (mb, 1) after the swap holds
// outputs 32 mb + 8 i + r and 32 mb + 8 i + 4 + r.  The swap is its own inverse: lane_to_rows
// takes per-point values (rows >= FO zero) back to the block layout.
__device__ __forceinline__ void rows_to_lane(const f32x16 (&C)[2][2], float (&o)[FO])
{
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int oa = 32 * mb + 8 * (v >> 2) + (v & 3), ob = oa + 4;
            if (oa < FO) {
                const Swapped t = swap32(__float_as_uint(C[mb][0][v]), __float_as_uint(C[mb][1][v]));
                o[oa] = __uint_as_float(t.a);
                if (ob < FO) o[ob] = __uint_as_float(t.b);
            }
        }
}
__device__ __forceinline__ void lane_to_rows(const float (&d)[FO], f32x16 (&C)[2][2])
{
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int oa = 32 * mb + 8 * (v >> 2) + (v & 3), ob = oa + 4;
            if (oa < FO) {
                const Swapped t = swap32(__float_as_uint(d[oa]), ob < FO ? __float_as_uint(d[ob]) : 0u);
                C[mb][0][v] = __uint_as_float(t.a);
                C[mb][1][v] = __uint_as_float(t.b);
            } else {
                C[mb][0][v] = 0.0f;
                C[mb][1][v] = 0.0f;
            }
        }
}

// Hardware transcendentals (about 1 ulp).
__device__ __forceinline__ float exp2_hw(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float log_hw(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }

// softplus(x) = max(x, 0) + log1p(e^-|x|), log1p(e) as log(u) e / (u - 1) with u = 1 + e (exact
// to a few ulp where u rounds; e itself where u == 1); also the logistic sigma(x) = softplus'(x).
__device__ __forceinline__ float softplus_hw(float x, float *sigma = nullptr)
{
    const float e = exp2_hw(-fabsf(x) * 1.4426950408889634f);
    const float u = 1.0f + e;
    const float l = u == 1.0f ? e : log_hw(u) * (e * rcp_hw(u - 1.0f));
    if (sigma) {
        const float iu = rcp_hw(u);
        *sigma = x >= 0.0f ? iu : e * iu;
    }
    return fmaxf(x, 0.0f) + l;
}

// Element i (0..2) of a 3-vector by selects, so that a per-launch index keeps the vector in registers.
__device__ __forceinline__ float pick3(const float *v, int i) { return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]); }
__device__ __forceinline__ void set3(float *v, int i, float x)
{
    v[0] = i == 0 ? x : v[0];
    v[1] = i == 1 ? x : v[1];
    v[2] = i == 2 ? x : v[2];
}
__device__ __forceinline__ int cond0(int d) { return d == 0 ? 1 : 0; }
__device__ __forceinline__ int cond1(int d) { return d == 2 ? 1 : 2; }

// The backward kernel over all layers (mi_flow_bwd.hip): g (per point) or, with g null, the
// variance-loss coefficient from the weights w and the per-block sums of w and w^2 in parts
// (record layout, nparts records); gradient rows of layer l at gp + (l * gp_stride + block) * FPL;
// hist (optional) receives the batch's (sum w, sum w^2).
hipError_t launch_flow_bwd(hipStream_t stream, uint32_t blocks, const FwdLayer *fl, const BwdExtra *bl, uint64_t n,
                           const float *xs, uint64_t xstride, const float *g, const float *w, const double *parts,
                           uint32_t nparts, float *abar, float *gp, uint32_t gp_stride, double *hist, int step);

}  // namespace mi
