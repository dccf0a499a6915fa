// lps_mlp.hip -- the learned PCISPH warm start: a per-particle MLP
// 16 -> 64 -> 64 -> 1 (ReLU) on BF16 matrix cores with FP32 accumulation.
//
// The network is evaluated in transposed form, Y = W X, with particles along
// the N dimension of v_mfma_f32_16x16x16_bf16: a 16x16 accumulator tile
// (lane l holds rows 4*(l/16)..+3 of column l%16) is exactly the B operand of
// the next layer (lane l holds K rows 4*(l/16)..+3 of column l%16), so the
// forward pass needs no data movement between layers.  The weight-gradient
// products reduce over particles and take their operands from LDS transposes.
//
// BF16: inputs (after normalisation), weights and hidden activations are
// rounded to BF16 (round to nearest even); biases, accumulators, the output
// layer's sum and all gradients accumulate in FP32.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>

#include "lps_common.h"
#include "lps_kernels.h"

using lps_dev::MlpArgs;
using lps_dev::RowMap;

namespace {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int IN = LPS_MLP_IN;       // 16
constexpr int HID = LPS_MLP_HIDDEN;  // 64
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + HID * IN;
constexpr int OFF_W2 = OFF_B1 + HID;
constexpr int OFF_B2 = OFF_W2 + HID * HID;
constexpr int OFF_W3 = OFF_B2 + HID;
constexpr int OFF_B3 = OFF_W3 + HID;
static_assert(OFF_B3 + 1 == LPS_MLP_PARAMS, "parameter layout");

// Round to nearest even bfloat16 (v_cvt_pk_bf16_f32; equal to rounding the bit
// pattern for every non-NaN float, a NaN stays a NaN: tests/bf16_check.hip).
__device__ __forceinline__ unsigned short f2bf(float f) { return __builtin_bit_cast(unsigned short, (__bf16)f); }

typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2v __attribute__((ext_vector_type(2)));
typedef unsigned u32x2v __attribute__((ext_vector_type(2)));

// Two roundings in one instruction: a in the low half, b in the high half.
__device__ __forceinline__ uint32_t f2bf2(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2v{a, b}, bf16x2v));
}
__device__ __forceinline__ bf16x4 f2bf4(float a, float b, float c, float d) {
    return __builtin_bit_cast(bf16x4, u32x2v{f2bf2(a, b), f2bf2(c, d)});
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float rbf(float f) { return bf2f(f2bf(f)); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x4 load4(const unsigned short* p) {
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFFu);
    r[1] = (short)(v.x >> 16);
    r[2] = (short)(v.y & 0xFFFFu);
    r[3] = (short)(v.y >> 16);
    return r;
}

struct alignas(16) MlpSmem {
    unsigned short w1[HID * IN];    // [neuron][feature]
    unsigned short w2[HID * HID];   // [out][in]
    unsigned short w2t[HID * HID];  // [in][out]
    float b1[HID], b2[HID], w3[HID];
    float mean[IN], istd[IN];
    float b3;
};

__device__ void load_weights(MlpSmem& S, const MlpArgs& m, bool transposed) {
    for (uint32_t k = threadIdx.x; k < HID * IN; k += blockDim.x) S.w1[k] = f2bf(m.params[OFF_W1 + k]);
    for (uint32_t k = threadIdx.x; k < HID * HID; k += blockDim.x) {
        const unsigned short w = f2bf(m.params[OFF_W2 + k]);
        S.w2[k] = w;
        if (transposed) S.w2t[(k % HID) * HID + k / HID] = w;
    }
    for (uint32_t k = threadIdx.x; k < HID; k += blockDim.x) {
        S.b1[k] = m.params[OFF_B1 + k];
        S.b2[k] = m.params[OFF_B2 + k];
        S.w3[k] = rbf(m.params[OFF_W3 + k]);
    }
    for (uint32_t k = threadIdx.x; k < IN; k += blockDim.x) {
        S.mean[k] = m.fscale[k];
        S.istd[k] = m.fscale[IN + k];
    }
    if (threadIdx.x == 0) S.b3 = m.params[OFF_B3];
}

// Forward pass of one group of 16 samples (column l%16 of the tiles).
struct Fwd {
    bf16x4 x;        // input tile (B layout): features 4*(l/16)..+3 of sample l%16
    f32x4 y1[4];     // layer-1 pre-activations, neuron tiles
    bf16x4 h1[4];
    f32x4 y2[4];
    bf16x4 h2[4];
    float y;         // output (identical in the four lane groups of a sample)
};

__device__ __forceinline__ void forward(const MlpSmem& S, const float* __restrict__ xrow, bool valid, Fwd& F) {
    const uint32_t l = threadIdx.x & 63u, g = l >> 4, r16 = l & 15u;
    for (int r = 0; r < 4; r++) {
        const uint32_t f = 4u * g + (uint32_t)r;
        const float v = valid ? (xrow[f] - S.mean[f]) * S.istd[f] : 0.f;
        F.x[r] = (short)f2bf(v);
    }
    for (int mt = 0; mt < 4; mt++) {
        const bf16x4 a = load4(&S.w1[(mt * 16 + r16) * IN + 4 * g]);
        F.y1[mt] = mfma(a, F.x, f32x4{0.f, 0.f, 0.f, 0.f});
        for (int r = 0; r < 4; r++) {
            const float v = F.y1[mt][r] + S.b1[mt * 16 + 4 * g + r];
            F.y1[mt][r] = v;
            F.h1[mt][r] = (short)f2bf(fmaxf(v, 0.f));
        }
    }
    for (int mt = 0; mt < 4; mt++) {
        f32x4 acc{0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < 4; kt++) acc = mfma(load4(&S.w2[(mt * 16 + r16) * HID + kt * 16 + 4 * g]), F.h1[kt], acc);
        for (int r = 0; r < 4; r++) {
            const float v = acc[r] + S.b2[mt * 16 + 4 * g + r];
            F.y2[mt][r] = v;
            F.h2[mt][r] = (short)f2bf(fmaxf(v, 0.f));
        }
    }
    float part = 0.f;
    for (int mt = 0; mt < 4; mt++)
        for (int r = 0; r < 4; r++) part += S.w3[mt * 16 + 4 * g + r] * bf2f((unsigned short)F.h2[mt][r]);
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    F.y = part + S.b3;
}

// The inference weights in BF16, packed once per change of the FP32 master
// copy (k_mlp_pack) so that the seed kernel copies them instead of converting
// them in every block.  The weight matrices are stored as MFMA A-operand
// fragments: fragment t holds, for lane l, the four values the lane passes
// to v_mfma_f32_16x16x16_bf16 (row 16 m + l%16, columns 16 k + 4 (l/16)..+3),
// at [t][l][4].  A wave's fragment read is then 512 contiguous bytes, free of
// LDS bank conflicts (a [row][column] layout puts the 16 rows a lane group
// reads 128 bytes apart, all in one bank).
struct alignas(16) MlpPacked {
    unsigned short w1[HID * IN];    // fragments (m), m < 4: W1[16 m + l%16][4 (l/16) + r]
    unsigned short w2[HID * HID];   // fragments (4 m + k): W2[16 m + l%16][16 k + 4 (l/16) + r]
    float b1[HID], b2[HID], w3[HID];
    float mean[IN], istd[IN];
    float b3;
    float pad_[3];
};
static_assert(sizeof(MlpPacked) % 16 == 0, "packed weights are copied as uint4");

// Inference of one group of 16 samples (column l%16), the output only: the
// arithmetic of forward() (same products, order and roundings) without
// keeping the pre-activations and hidden tiles the gradient needs.
// xin holds features 4*(l/16)..+3 of the lane's sample.
__device__ __forceinline__ float infer(const MlpPacked& S, const float4& xin, bool valid) {
    const uint32_t l = threadIdx.x & 63u, g = l >> 4;
    const float xv[4] = {xin.x, xin.y, xin.z, xin.w};
    float v[4];
    for (int r = 0; r < 4; r++) {
        const uint32_t f = 4u * g + (uint32_t)r;
        v[r] = valid ? (xv[r] - S.mean[f]) * S.istd[f] : 0.f;
    }
    const bf16x4 x = f2bf4(v[0], v[1], v[2], v[3]);
    bf16x4 h1[4];
    for (int mt = 0; mt < 4; mt++) {
        const f32x4 y = mfma(load4(&S.w1[(mt * 64 + l) * 4]), x, f32x4{0.f, 0.f, 0.f, 0.f});
        const float* b1 = &S.b1[mt * 16 + 4 * g];
        h1[mt] = f2bf4(fmaxf(y[0] + b1[0], 0.f), fmaxf(y[1] + b1[1], 0.f), fmaxf(y[2] + b1[2], 0.f),
                       fmaxf(y[3] + b1[3], 0.f));
    }
    float part = 0.f;
    for (int mt = 0; mt < 4; mt++) {
        f32x4 acc{0.f, 0.f, 0.f, 0.f};
        for (int kt = 0; kt < 4; kt++) acc = mfma(load4(&S.w2[((mt * 4 + kt) * 64 + l) * 4]), h1[kt], acc);
        const float* b2 = &S.b2[mt * 16 + 4 * g];
        const float* w3 = &S.w3[mt * 16 + 4 * g];
        const uint32_t lo = f2bf2(fmaxf(acc[0] + b2[0], 0.f), fmaxf(acc[1] + b2[1], 0.f));
        const uint32_t hi = f2bf2(fmaxf(acc[2] + b2[2], 0.f), fmaxf(acc[3] + b2[3], 0.f));
        part += w3[0] * __uint_as_float(lo << 16);
        part += w3[1] * __uint_as_float(lo & 0xFFFF0000u);
        part += w3[2] * __uint_as_float(hi << 16);
        part += w3[3] * __uint_as_float(hi & 0xFFFF0000u);
    }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    return part + S.b3;
}

__global__ __launch_bounds__(LPS_BLOCK) void k_mlp_pack(const float* __restrict__ params,
                                                        const float* __restrict__ fscale, MlpPacked* __restrict__ out) {
    // element e of the fragment array: r = e % 4, lane l = (e / 4) % 64,
    // fragment t = e / 256
    for (uint32_t e = threadIdx.x; e < HID * IN; e += blockDim.x) {
        const uint32_t r = e & 3u, l = (e >> 2) & 63u, m = e >> 8;
        out->w1[e] = f2bf(params[OFF_W1 + (16u * m + (l & 15u)) * IN + 4u * (l >> 4) + r]);
    }
    for (uint32_t e = threadIdx.x; e < HID * HID; e += blockDim.x) {
        const uint32_t r = e & 3u, l = (e >> 2) & 63u, t = e >> 8, m = t >> 2, k = t & 3u;
        out->w2[e] = f2bf(params[OFF_W2 + (16u * m + (l & 15u)) * HID + 16u * k + 4u * (l >> 4) + r]);
    }
    for (uint32_t k = threadIdx.x; k < HID; k += blockDim.x) {
        out->b1[k] = params[OFF_B1 + k];
        out->b2[k] = params[OFF_B2 + k];
        out->w3[k] = rbf(params[OFF_W3 + k]);
    }
    for (uint32_t k = threadIdx.x; k < IN; k += blockDim.x) {
        out->mean[k] = fscale[k];
        out->istd[k] = fscale[IN + k];
    }
    if (threadIdx.x == 0) {
        out->b3 = params[OFF_B3];
        out->pad_[0] = out->pad_[1] = out->pad_[2] = 0.f;
    }
}

// Copy of the packed weights into LDS by a block of LPS_BLOCK threads: every
// load is issued before the first store, so the copy waits for one memory
// round trip rather than one per row of the block.
__device__ __forceinline__ void load_packed(MlpPacked& S, const MlpPacked* __restrict__ packed) {
    constexpr uint32_t N = sizeof(MlpPacked) / 16u;
    static_assert(N > 2u * LPS_BLOCK && N <= 3u * LPS_BLOCK, "three rows of uint4 per thread");
    const uint4* src = reinterpret_cast<const uint4*>(packed);
    uint4* dst = reinterpret_cast<uint4*>(&S);
    const uint32_t t = threadIdx.x;
    const bool third = t + 2u * LPS_BLOCK < N;
    const uint4 a = src[t], b = src[t + LPS_BLOCK];
    const uint4 c = third ? src[t + 2u * LPS_BLOCK] : make_uint4(0u, 0u, 0u, 0u);
    dst[t] = a;
    dst[t + LPS_BLOCK] = b;
    if (third) dst[t + 2u * LPS_BLOCK] = c;
}

// Network output (in units of the target scale) of n rows of x.
__global__ __launch_bounds__(LPS_BLOCK) void k_mlp_forward(const MlpPacked* __restrict__ packed,
                                                           const float* __restrict__ x, uint32_t n,
                                                           float* __restrict__ y) {
    __shared__ MlpPacked S;
    load_packed(S, packed);
    __syncthreads();
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    for (uint32_t bb = blockIdx.x * LPS_BLOCK; bb < n; bb += gridDim.x * LPS_BLOCK)
        for (uint32_t grp = 0; grp < 4; grp++) {
            const uint32_t p = bb + w * 64u + grp * 16u + (l & 15u);
            const float4 xin = *reinterpret_cast<const float4*>(x + (size_t)min(p, n - 1u) * IN + 4u * (l >> 4));
            const float v = infer(S, xin, p < n);
            if (p < n && (l >> 4) == 0u) y[p] = v;
        }
}

}  // namespace

// Seed of the PCISPH pressure: p_i = pscale * tscale * MLP(features_i) with
// pscale = delta * rho0 (the network works in units of delta * rho0, which
// makes it independent of particle size and time step), clamped at zero when
// negative pressures are disabled, non-finite values replaced by 0;
// q_i = p_i / rho*_i^2 for the pressure-force pass.
// Four waves per SIMD: the compiler keeps the layer weights and biases in
// registers across the four sample groups (128 VGPRs, no spills; five waves
// would spill).
__global__ __launch_bounds__(LPS_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void k_mlp_seed(const MlpPacked* __restrict__ packed, float tscale,
                                                        const float* __restrict__ feat, uint32_t fs, uint32_t n,
                                                        float* __restrict__ press, float4* __restrict__ xq,
                                                        float* __restrict__ seed, const float* __restrict__ rho_pred,
                                                        const float* __restrict__ pprev, float kappa,
                                                        float pscale, int clamp_nonneg, const DevGrid* g,
                                                        lps_dev::Scenes sc, LoopRule rule, float rest_density) {
    if (g->status != 0u) return;
    __shared__ MlpPacked S;
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u;
    // one block per 256-slot chunk, which is one particle block and so belongs
    // to one scene; chunks of scenes that stopped at the first correction are
    // skipped.  The wave's four groups of 16 samples load their features
    // first, so that their latency overlaps the copy of the weights.
    const uint32_t bb = blockIdx.x * LPS_BLOCK;
    // features are feature-major (feature f of slot p at f * fs + p)
    float4 xin[4];
    for (uint32_t grp = 0; grp < 4; grp++) {
        const uint32_t p = min(bb + w * 64u + grp * 16u + (l & 15u), n - 1u);
        const float* fp = feat + (size_t)(4u * (l >> 4)) * fs + p;
        xin[grp] = make_float4(fp[0], fp[fs], fp[2 * (size_t)fs], fp[3 * (size_t)fs]);
    }
    load_packed(S, packed);
    {
        // the seed replaces the first correction's pressure unless the scene's
        // loop stops there (the loop rule at k = 1, which the pressure pass
        // applies: a non-finite error, the tolerance met with min <= 1, or max
        // <= 1; the stall rule cannot stop a loop at its first correction).
        // The test is uniform over the block, so the whole block leaves
        // before the barrier or none of it does.
        const uint32_t s = sc.block_scene[blockIdx.x];
        const uint32_t eb = sc.grid[s].err_bits[0];
        if (eb >= 0x7F800000u || (1u >= rule.min_iter && __uint_as_float(eb) / rest_density < rule.tol) ||
            1u >= rule.max_iter)
            return;
        const uint32_t end = min(n, sc.info[s].slot0 + sc.info[s].count);
        // the first correction's pressure buffer of this scene
        const size_t ob = (size_t)sc.grid[s].buf_of[1] * fs;
        __syncthreads();
        for (uint32_t grp = 0; grp < 4; grp++) {
            const uint32_t p = bb + w * 64u + grp * 16u + (l & 15u);
            const float y = infer(S, xin[grp], p < end);
            if (p < end && (l >> 4) == 0u) {
                float v = y * tscale * pscale;
                if (!isfinite(v)) v = 0.f;
                seed[p] = v;
                if (clamp_nonneg) {
                    // With negative pressures disabled the loop never lowers
                    // a pressure, so the seed stays within an envelope: at
                    // least the zero start's own first-correction pressure
                    // p1 (which the loop would reach anyway), at most kappa
                    // times the particle's previous applied pressure unless
                    // p1 is larger.  An over-prediction then shrinks by kappa
                    // per step instead of feeding itself through the
                    // previous-pressure feature.
                    v = fmaxf(press[ob + p], fminf(v, kappa * pprev[p]));
                }
                const float rp = rho_pred[ob + p];
                press[ob + p] = v;
                // the pressure pass reads p / rho*^2 from the w slot of the prediction copy
                reinterpret_cast<float*>(xq)[4u * p + 3u] = v / (rp * rp);
            }
        }
    }
}


// p[0..n) = 0 (clears accumulators without a memset node)
__global__ void k_fill_zero(float* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}

// Gradient of the weighted mean squared error mean(w d^2), d = y - t/tscale,
// w = over_weight where d > 0 and 1 elsewhere, over a batch of n rows of x / t
// (row of sample s given by rm) with respect to the FP32 parameters, added to
// grad; loss gets the batch's loss.  The
// forward uses the BF16-rounded values (straight-through).  Each block sums
// its samples' gradients in LDS, then adds them to grad with global atomics.
__global__ __launch_bounds__(LPS_BLOCK) void k_mlp_grad(MlpArgs m, const float* __restrict__ x,
                                                        const float* __restrict__ t, RowMap rm, uint32_t n,
                                                        float inv_n, float over_weight, float* __restrict__ grad,
                                                        float* __restrict__ loss) {
    __shared__ MlpSmem S;
    // load4 reads 8 bytes at a time: every staging array is 16-byte aligned
    __shared__ __align__(16) float G[LPS_MLP_PARAMS];
    __shared__ __align__(16) unsigned short T_[LPS_BLOCK / 64][HID * 16];   // dY2 [neuron][sample]
    __shared__ __align__(16) unsigned short U_[LPS_BLOCK / 64][HID * 16];   // H1  [neuron][sample]
    __shared__ __align__(16) unsigned short D_[LPS_BLOCK / 64][HID * 16];   // dY1 [neuron][sample]
    __shared__ __align__(16) unsigned short V_[LPS_BLOCK / 64][IN * 16];    // X   [feature][sample]
    __shared__ float s_loss;
    load_weights(S, m, true);
    for (uint32_t k = threadIdx.x; k < LPS_MLP_PARAMS; k += blockDim.x) G[k] = 0.f;
    if (threadIdx.x == 0) s_loss = 0.f;
    __syncthreads();
    const uint32_t w = threadIdx.x >> 6, l = threadIdx.x & 63u, g = l >> 4, r16 = l & 15u;
    unsigned short* T = T_[w];
    unsigned short* U = U_[w];
    unsigned short* D = D_[w];
    unsigned short* V = V_[w];
    f32x4 dw2[4][4], dw1[4];
    for (int a = 0; a < 4; a++) {
        dw1[a] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int b = 0; b < 4; b++) dw2[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    float gw3[16], gb2[16], gb1[16];
    for (int k = 0; k < 16; k++) gw3[k] = gb2[k] = gb1[k] = 0.f;
    float gb3 = 0.f, lsum = 0.f;
    // block-uniform trip count: the loop body holds barriers
    for (uint32_t bb = blockIdx.x * LPS_BLOCK; bb < n; bb += gridDim.x * LPS_BLOCK)
    for (uint32_t grp = 0; grp < 4; grp++) {
        const uint32_t s = bb + w * 64u + grp * 16u + r16;
        const bool valid = s < n;
        const uint32_t row = valid ? rm.row(s) : 0u;
        Fwd F;
        forward(S, x + (size_t)row * IN, valid, F);
        const float tn = valid ? t[row] / m.tscale : 0.f;
        const float d = valid ? F.y - tn : 0.f;
        const float wd = d > 0.f ? over_weight : 1.f;  // over-predictions weigh more
        const float dy = 2.f * wd * d * inv_n;
        if (g == 0u) lsum += wd * d * d;
        // output layer
        if (g == 0u) gb3 += dy;
        bf16x4 dy2b[4];
        for (int mt = 0; mt < 4; mt++)
            for (int r = 0; r < 4; r++) {
                const int nn = mt * 16 + 4 * (int)g + r;
                gw3[mt * 4 + r] += dy * bf2f((unsigned short)F.h2[mt][r]);
                const float dy2 = F.y2[mt][r] > 0.f ? S.w3[nn] * dy : 0.f;
                gb2[mt * 4 + r] += dy2;
                dy2b[mt][r] = (short)f2bf(dy2);
            }
        // dH1 = W2^T dY2, dY1 = dH1 * relu'(Y1)
        bf16x4 dy1b[4];
        for (int mt = 0; mt < 4; mt++) {
            f32x4 acc{0.f, 0.f, 0.f, 0.f};
            for (int kt = 0; kt < 4; kt++)
                acc = mfma(load4(&S.w2t[(mt * 16 + r16) * HID + kt * 16 + 4 * g]), dy2b[kt], acc);
            for (int r = 0; r < 4; r++) {
                const float dy1 = F.y1[mt][r] > 0.f ? acc[r] : 0.f;
                gb1[mt * 4 + r] += dy1;
                dy1b[mt][r] = (short)f2bf(dy1);
            }
        }
        // stage [neuron][sample] tiles for the weight gradients
        for (int mt = 0; mt < 4; mt++)
            for (int r = 0; r < 4; r++) {
                const int nn = mt * 16 + 4 * (int)g + r;
                T[nn * 16 + r16] = (unsigned short)dy2b[mt][r];
                U[nn * 16 + r16] = (unsigned short)F.h1[mt][r];
                D[nn * 16 + r16] = (unsigned short)dy1b[mt][r];
            }
        for (int r = 0; r < 4; r++) V[(4 * g + r) * 16 + r16] = (unsigned short)F.x[r];
        __syncthreads();
        // dW2 += dY2 H1^T  (K = samples); A[o][s] = T, B[s][i] = U^T
        for (int mt = 0; mt < 4; mt++) {
            const bf16x4 a = load4(&T[(mt * 16 + r16) * 16 + 4 * g]);
            for (int nt = 0; nt < 4; nt++) dw2[mt][nt] = mfma(a, load4(&U[(nt * 16 + r16) * 16 + 4 * g]), dw2[mt][nt]);
        }
        // dW1 += dY1 X^T; A[n][s] = D, B[s][f] = V^T
        {
            const bf16x4 b = load4(&V[r16 * 16 + 4 * g]);
            for (int mt = 0; mt < 4; mt++) dw1[mt] = mfma(load4(&D[(mt * 16 + r16) * 16 + 4 * g]), b, dw1[mt]);
        }
        __syncthreads();
    }
    // reduce per-sample sums over the 16 samples of a lane group, then into LDS
    for (int k = 0; k < 16; k++) {
        float a = gw3[k], b = gb2[k], c = gb1[k];
        for (int off = 1; off < 16; off <<= 1) {
            a += __shfl_xor(a, off, 64);
            b += __shfl_xor(b, off, 64);
            c += __shfl_xor(c, off, 64);
        }
        if (r16 == 0u) {
            const int nn = (k / 4) * 16 + 4 * (int)g + (k % 4);
            atomicAdd(&G[OFF_W3 + nn], a);
            atomicAdd(&G[OFF_B2 + nn], b);
            atomicAdd(&G[OFF_B1 + nn], c);
        }
    }
    for (int off = 1; off < 16; off <<= 1) gb3 += __shfl_xor(gb3, off, 64);
    if (l == 0u) atomicAdd(&G[OFF_B3], gb3);
    for (int off = 1; off < 64; off <<= 1) lsum += __shfl_xor(lsum, off, 64);
    if (l == 0u) atomicAdd(&s_loss, lsum);
    for (int mt = 0; mt < 4; mt++)
        for (int r = 0; r < 4; r++) {
            const int o = mt * 16 + 4 * (int)g + r;
            for (int nt = 0; nt < 4; nt++) atomicAdd(&G[OFF_W2 + o * HID + nt * 16 + (int)r16], dw2[mt][nt][r]);
            atomicAdd(&G[OFF_W1 + o * IN + (int)r16], dw1[mt][r]);
        }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < LPS_MLP_PARAMS; k += blockDim.x)
        if (G[k] != 0.f) atomicAdd(&grad[k], G[k]);
    if (threadIdx.x == 0 && loss) atomicAdd(loss, s_loss * inv_n);
}

// Adam on the FP32 master parameters (bias-corrected); the gradient buffer is
// cleared for the next batch.
__global__ void k_adam(float* __restrict__ params, float* __restrict__ grad, float* __restrict__ m1,
                       float* __restrict__ m2, uint32_t count, float lr, float b1, float b2, float eps, float c1,
                       float c2) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const float gk = grad[k];
    const float a = b1 * m1[k] + (1.f - b1) * gk;
    const float b = b2 * m2[k] + (1.f - b2) * gk * gk;
    m1[k] = a;
    m2[k] = b;
    params[k] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    grad[k] = 0.f;
}

// Per-column sums of x and x^2 (double) over n rows of `cols` (<= 16)
// columns, added to acc[0..cols) and acc[cols..2 cols).
__global__ __launch_bounds__(LPS_BLOCK) void k_feature_stats(const float* __restrict__ x, uint64_t n, uint32_t cols,
                                                             double* __restrict__ acc) {
    __shared__ double s[2 * IN];
    if (threadIdx.x < 2 * IN) s[threadIdx.x] = 0.0;
    __syncthreads();
    double a[IN], b[IN];
    for (int f = 0; f < IN; f++) a[f] = b[f] = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        for (int f = 0; f < IN; f++)
            if ((uint32_t)f < cols) {
                const double v = x[i * cols + f];
                a[f] += v;
                b[f] += v * v;
            }
    for (int f = 0; f < IN; f++)
        if ((uint32_t)f < cols) {
            atomicAdd(&s[f], a[f]);
            atomicAdd(&s[IN + f], b[f]);
        }
    __syncthreads();
    if (threadIdx.x < cols) {
        atomicAdd(&acc[threadIdx.x], s[threadIdx.x]);
        atomicAdd(&acc[cols + threadIdx.x], s[IN + threadIdx.x]);
    }
}

// Training-set append: particle offset + r * stride (r < count) gives row
// base + r: its features (feature-major source, feature f at f * fs + i) and
// its final pressure divided by pscale.
__global__ void k_collect(const float* __restrict__ feat, uint32_t fs, const float* __restrict__ press,
                          float inv_pscale, uint32_t n, uint32_t stride, uint32_t offset, uint32_t count,
                          float* __restrict__ X, float* __restrict__ Tg, uint64_t base) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    const uint64_t i = (uint64_t)offset + (uint64_t)r * stride;
    if (i >= n) return;
    float* dst = X + (base + r) * IN;
    for (int f = 0; f < IN; f++) dst[f] = feat[(uint64_t)f * fs + i];
    Tg[base + r] = press[i] * inv_pscale;
}

namespace lps_dev {

namespace {
uint32_t mlp_grid(uint64_t n) {
    const uint64_t nb = (n + LPS_BLOCK - 1u) / LPS_BLOCK;
    return (uint32_t)(nb < 1024u ? (nb ? nb : 1u) : 1024u);
}
}  // namespace

size_t mlp_packed_bytes() { return sizeof(MlpPacked); }

void launch_mlp_pack(hipStream_t s, const MlpArgs& m, void* packed) {
    k_mlp_pack<<<1, LPS_BLOCK, 0, s>>>(m.params, m.fscale, static_cast<MlpPacked*>(packed));
}

void launch_mlp_seed(hipStream_t s, const void* packed, float tscale, const float* feat, uint32_t fs, uint32_t n,
                     float* press, float4* xq, float* seed, const float* rho_pred, const float* pprev,
                     float kappa, float pscale, int clamp_nonneg, const DevGrid* g, const Scenes& sc,
                     const LoopRule& rule, float rest_density) {
    // one block per particle block (n is a whole number of blocks)
    k_mlp_seed<<<(n + LPS_BLOCK - 1u) / LPS_BLOCK, LPS_BLOCK, 0, s>>>(static_cast<const MlpPacked*>(packed), tscale,
                                                                      feat, fs, n, press, xq, seed, rho_pred, pprev,
                                                                      kappa, pscale, clamp_nonneg, g, sc, rule,
                                                                      rest_density);
}

void launch_fill_zero(hipStream_t s, float* p, uint32_t n) {
    if (n == 0) return;
    k_fill_zero<<<(n + 255u) / 256u, 256, 0, s>>>(p, n);
}

void launch_mlp_forward(hipStream_t s, const void* packed, const float* x, uint32_t n, float* y) {
    k_mlp_forward<<<mlp_grid(n), LPS_BLOCK, 0, s>>>(static_cast<const MlpPacked*>(packed), x, n, y);
}

void launch_mlp_grad(hipStream_t s, const MlpArgs& m, const float* x, const float* t, const RowMap& rm, uint32_t n,
                     float over_weight, float* grad, float* loss) {
    k_mlp_grad<<<mlp_grid(n), LPS_BLOCK, 0, s>>>(m, x, t, rm, n, 1.f / (float)n, over_weight, grad, loss);
}

void launch_adam(hipStream_t s, float* params, float* grad, float* m1, float* m2, uint32_t count, float lr, float b1,
                 float b2, float eps, uint32_t step) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    k_adam<<<(count + 255u) / 256u, 256, 0, s>>>(params, grad, m1, m2, count, lr, b1, b2, eps, c1, c2);
}

void launch_feature_stats(hipStream_t s, const float* x, uint64_t n, uint32_t cols, double* acc) {
    k_feature_stats<<<256, LPS_BLOCK, 0, s>>>(x, n, cols, acc);
}

void launch_collect(hipStream_t s, const float* feat, uint32_t fs, const float* press, float pscale, uint32_t n,
                    uint32_t stride, uint32_t offset, uint32_t count, float* X, float* T, uint64_t base) {
    if (count == 0) return;
    k_collect<<<(count + 255u) / 256u, 256, 0, s>>>(feat, fs, press, 1.f / pscale, n, stride, offset, count, X, T,
                                                    base);
}

}  // namespace lps_dev
