// fdmamr_selector.hip -- learned troubled-cell selector (no reference counterpart).
//
// A 4-layer perceptron (75 -> 64 -> 64 -> 64 -> 1, GELU, sigmoid) reads each interior leaf's
// 5x5 stencil of (rho, p, |u|), each channel centred on the leaf and scaled by its largest
// deviation on the stencil (plus 1e-3 of |centre|, so flat data maps to ~0), and predicts
// whether the leaf's nonlinear WCNS3 weights deviate from the linear ones.
//
// Matrix cores: every layer is a BF16 MFMA (v_mfma_f32_16x16x16_bf16, FP32 accumulate)
// computed transposed, H_next^T = W^T H^T, so the 16x16 accumulator tile of one layer is
// exactly the B operand of the next (no LDS transposes on the forward path). A wave handles
// 16 leaves per tile; weights live in LDS as BF16, transposed for the A operand.
//
// Training (forward + backward) uses the same MFMA tiles; weight gradients are MFMA products
// with the leaf dimension as K, operands staged through LDS in [unit][leaf] layout.
// FP32 master weights, Adam, BF16 copies repacked after every update.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fdmamr_device.hpp"
#include "fdmamr_selector.hpp"

namespace fdm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// GELU and its derivative. g(x) = x s(c), c = x clamped to [-GELU_A, GELU_A], with
// s(c) = 1/2 + c P(c^2) and P of degree 5: the minimax fit (weighted by |x|, with s(GELU_A) = 1)
// of the tanh form's factor 1 / (1 + exp(-2 k0 (x + k1 x^3))), k0 = sqrt(2/pi), k1 = 0.044715.
// |g - GELU_tanh| <= 6.2e-4 and |g - x Phi(x)| <= 8.2e-4 everywhere, below the BF16 rounding
// of the hidden activations (2^-9 relative); g(x) = x for x >= GELU_A and 0 for x <= -GELU_A
// (up to the coefficients' rounding); g(x) - g(-x) = x. Nine full-rate VALU operations per
// unit and no transcendentals (the activation is the selector's largest VALU cost per leaf,
// 192 evaluations). g'(x) = s(c) + x s'(c) inside the clamp and s(c) outside.
constexpr float GELU_A = 3.5f;
constexpr float GELU_P0 = 0x1.958016p-2f, GELU_P1 = -0x1.fed952p-5f, GELU_P2 = 0x1.013ab0p-7f,
                GELU_P3 = -0x1.496018p-11f, GELU_P4 = 0x1.da140ap-16f, GELU_P5 = -0x1.215ff2p-21f;
__device__ __forceinline__ float gelu_clamp(float x) { return __builtin_amdgcn_fmed3f(x, -GELU_A, GELU_A); }
__device__ __forceinline__ float gelu_poly(float t) {
    float p = __builtin_fmaf(GELU_P5, t, GELU_P4);
    p = __builtin_fmaf(p, t, GELU_P3);
    p = __builtin_fmaf(p, t, GELU_P2);
    p = __builtin_fmaf(p, t, GELU_P1);
    return __builtin_fmaf(p, t, GELU_P0);
}
__device__ __forceinline__ float gelu(float x) {
    const float c = gelu_clamp(x);
    return x * __builtin_fmaf(c, gelu_poly(c * c), 0.5f);
}
__device__ __forceinline__ float gelu_d(float x) {
    const float c = gelu_clamp(x), t = c * c;
    const float s = __builtin_fmaf(c, gelu_poly(t), 0.5f);
    if (c != x) return s;
    // s'(c) = P(t) + 2 t P'(t)
    float q = 5.f * GELU_P5;
    q = __builtin_fmaf(q, t, 4.f * GELU_P4);
    q = __builtin_fmaf(q, t, 3.f * GELU_P3);
    q = __builtin_fmaf(q, t, 2.f * GELU_P2);
    q = __builtin_fmaf(q, t, GELU_P1);
    const float ds = __builtin_fmaf(2.f * t, q, gelu_poly(t));
    return __builtin_fmaf(x, ds, s);
}

__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a); r[1] = (short)f2bf(b); r[2] = (short)f2bf(c); r[3] = (short)f2bf(d);
    return r;
}
__device__ __forceinline__ bf16x4 lds4(const uint16_t* p) {
    bf16x4 r;
    r[0] = (short)p[0]; r[1] = (short)p[1]; r[2] = (short)p[2]; r[3] = (short)p[3];
    return r;
}

// Cross-row combinations of the four lanes l, l ^ 16, l ^ 32, l ^ 48 with gfx950's lane swaps
// instead of LDS permutes: swapping a register with itself across rows leaves the pair
// {value of l, value of l ^ 16} (or l ^ 32) in the two outputs of every lane, and a
// commutative combination of the pair is the same in both lanes.
__device__ __forceinline__ float sum4_rows(float v) {
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}
// the same for a maximum of non-negative floats (integer max of the bit patterns)
__device__ __forceinline__ float max4_rows_nonneg(float v) {
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    const int m = max((int)a[0], (int)a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap((unsigned)m, (unsigned)m, false, false);
    return __int_as_float(max((int)b[0], (int)b[1]));
}

// ---------------------------------------------------------------- features
// Index-table row of point p of the 5x5 stencil: 0..7 the leaf's own stencil rows (axis
// points), 8..23 the 16 off-axis rows, 24 the centre (and p > 24, which has no point).
__device__ constexpr int point_row(int p) {
    return p >= 25 ? 24
         : (p % 5 == 2 && p / 5 == 2) ? 24
         : (p / 5 == 2) ? (p % 5 == 3 ? 0 : p % 5 == 1 ? 1 : p % 5 == 4 ? 4 : 5)
         : (p % 5 == 2) ? (p / 5 == 3 ? 2 : p / 5 == 1 ? 3 : p / 5 == 4 ? 6 : 7)
         : 8 + (p / 5 - (p / 5 > 2 ? 1 : 0)) * 4 + (p % 5 - (p % 5 > 2 ? 1 : 0));
}

// State indices of the lane's 7 stencil points. Branch-free: every lane issues all 7 loads
// (the centre, points past the grid and leaves past N read an in-range entry that is then
// replaced by the leaf itself), so the loads of a tile travel together.
__device__ __forceinline__ void gather_idx(const DevTables& t, const int32_t* __restrict__ off16, int leaf,
                                           bool valid, int idx[7], bool no_hoist = false) {
    int g = (threadIdx.x & 63) >> 4;
    // in a loop's rare branch: the row selections are recomputed there instead of being hoisted
    // out of the loop (where their registers would be live across the whole loop)
    if (no_hoist) asm volatile("" : "+v"(g));
    const int lc = valid ? leaf : 0;
#pragma unroll
    for (int k = 0; k < 7; k++) {
        const int r0 = point_row(4 * k), r1 = point_row(4 * k + 1), r2 = point_row(4 * k + 2),
                  r3 = point_row(4 * k + 3);
        const int row = g == 0 ? r0 : g == 1 ? r1 : g == 2 ? r2 : r3;
        const int rr = row < 24 ? row : 0;
        const int raw = rr < 8 ? t.stencil[(size_t)lc * 8 + rr] : off16[(size_t)(rr - 8) * t.N + lc];
        idx[k] = row < 24 ? raw : lc;
    }
}

// Lane l of a wave computes for leaf (l % 16) of the tile the stencil points p = g, g+4, ...
// (g = l / 16) of the 5x5 grid (p = (b+2)*5 + (a+2)); the four lanes of a leaf combine their
// maxima by shuffles. Channel f of point p becomes (v - c) / (m + 1e-3 |c|), c the centre
// value and m the larger of |c| and the largest |v - c| on the stencil, so every feature lies
// in [-1, 1] and smooth data maps near 0. v[k][f] holds point g + 4k (0 where p > 24).
// Returns true when the leaf's stencil is near-uniform: in every channel the largest |v - c|
// is at most UNIFORM_REL |c| (exactly uniform data included).
constexpr float UNIFORM_REL = 1e-6f;

// The channels (rho, p, |u|) of an FP32 state (fourth component unused).
__device__ __forceinline__ float4 channels(const float4& q) {
    return make_float4(q.x, q.w, __builtin_amdgcn_sqrtf(__builtin_fmaf(q.y, q.y, q.z * q.z)), 0.f);
}

// cc: channels of the centre, q: channels of the lane's 7 points. The index tables give points
// past the grid (p > 24) the centre itself, so their features are exactly 0; a lane past N reads
// in-range entries, and its features (discarded) only enter its own leaf's network column.
__device__ __forceinline__ bool features_from(const float4& cc, const float4* q, float v[7][3]) {
    const float c[3] = {cc.x, cc.y, cc.z};
#pragma unroll
    for (int k = 0; k < 7; k++) {
        v[k][0] = q[k].x;
        v[k][1] = q[k].y;
        v[k][2] = q[k].z;
    }
    bool uniform = true;
#pragma unroll
    for (int f = 0; f < 3; f++) {
        float md = 0.f;
#pragma unroll
        for (int k = 0; k < 7; k++) md = fmaxf(md, fabsf(v[k][f] - c[f]));
        md = max4_rows_nonneg(md);
        uniform = uniform && md <= UNIFORM_REL * fabsf(c[f]);
        const float sc = __builtin_amdgcn_rcpf(fmaxf(md, fabsf(c[f])) + 1e-3f * fabsf(c[f]) + 1e-30f);
#pragma unroll
        for (int k = 0; k < 7; k++) v[k][f] = (v[k][f] - c[f]) * sc;
    }
    return uniform;
}

__device__ __forceinline__ bool gather_values(const float4* __restrict__ Qf, int leaf, bool valid,
                                              const int idx[7], float v[7][3]) {
    const float4 qc = channels(Qf[valid ? leaf : 0]);
    float4 q[7];
#pragma unroll
    for (int k = 0; k < 7; k++) q[k] = channels(Qf[idx[k]]);
    return features_from(qc, q, v);
}

__device__ __forceinline__ void gather_features(const DevTables& t, const float4* __restrict__ Qf,
                                                const int32_t* __restrict__ off16, int leaf, bool valid,
                                                float v[7][3]) {
    int idx[7];
    gather_idx(t, off16, leaf, valid, idx);
    (void)gather_values(Qf, leaf, valid, idx, v);
}

// Writes the 80 BF16 features of the lane's leaf (75 used, feature p*3 + f) to xrow.
__device__ __forceinline__ void leaf_features(const DevTables& t, const float4* __restrict__ Qf,
                                              const int32_t* __restrict__ off16, int leaf, bool valid,
                                              uint16_t* xrow) {
    const int g = (threadIdx.x & 63) >> 4;
    float v[7][3];
    gather_features(t, Qf, off16, leaf, valid, v);
#pragma unroll
    for (int k = 0; k < 7; k++) {
        const int p = g + 4 * k;
        if (p < 25)
#pragma unroll
            for (int f = 0; f < 3; f++) xrow[p * 3 + f] = f2bf(v[k][f]);
    }
    if (g == 0)
#pragma unroll
        for (int f = 75; f < MLP_IN; f++) xrow[f] = 0;
}

// ---------------------------------------------------------------- forward (one 16-leaf tile)
struct FwdOut {
    f32x4 z[3][4];     // pre-activations per layer, 4 unit tiles (C layout: unit 16mt+4(l/16)+j, leaf l%16)
    f32x4 h3[4];       // last hidden layer (FP32)
    float z4;          // logit of leaf l%16 (valid in every lane of the leaf)
};

// sW: LDS weights (W1T [64][80], W2T [64][64], W3T [64][64] as BF16); sB: b1,b2,b3,w4 (FP32) + b4
__device__ __forceinline__ void mlp_forward(const uint16_t* sW, const float* sB, const uint16_t* xt,
                                            FwdOut& o, uint16_t* sH1, uint16_t* sH2) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    bf16x4 hb[4];
    // layer 1: K = 80 (5 k-steps); B operand from the X tile [leaf][80]
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        f32x4 acc;
#pragma unroll
        for (int j = 0; j < 4; j++) acc[j] = sB[SB_B1 + 16 * mt + 4 * g + j];
#pragma unroll
        for (int s = 0; s < 5; s++) {
            const bf16x4 a = lds4(sW + KW_W1T + (16 * mt + r) * MLP_IN + 16 * s + 4 * g);
            const bf16x4 b = lds4(xt + r * MLP_IN + 16 * s + 4 * g);
            acc = mfma16(a, b, acc);
        }
        o.z[0][mt] = acc;
    }
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        float h[4];
#pragma unroll
        for (int j = 0; j < 4; j++) h[j] = gelu(o.z[0][mt][j]);
        hb[mt] = pack4(h[0], h[1], h[2], h[3]);
        if (sH1)
#pragma unroll
            for (int j = 0; j < 4; j++) sH1[(16 * mt + 4 * g + j) * 16 + r] = (uint16_t)hb[mt][j];
    }
    // layers 2 and 3: K = 64 (4 k-steps); B operand = previous accumulator tiles
#pragma unroll
    for (int L = 1; L < 3; L++) {
        const int woff = L == 1 ? KW_W2T : KW_W3T;
        const int boff = L == 1 ? SB_B2 : SB_B3;
#pragma unroll
        for (int mt = 0; mt < 4; mt++) {
            f32x4 acc;
#pragma unroll
            for (int j = 0; j < 4; j++) acc[j] = sB[boff + 16 * mt + 4 * g + j];
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const bf16x4 a = lds4(sW + woff + (16 * mt + r) * MLP_H + 16 * s + 4 * g);
                acc = mfma16(a, hb[s], acc);
            }
            o.z[L][mt] = acc;
        }
        if (L == 1) {
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                float h[4];
#pragma unroll
                for (int j = 0; j < 4; j++) h[j] = gelu(o.z[1][mt][j]);
                hb[mt] = pack4(h[0], h[1], h[2], h[3]);
                if (sH2)
#pragma unroll
                    for (int j = 0; j < 4; j++) sH2[(16 * mt + 4 * g + j) * 16 + r] = (uint16_t)hb[mt][j];
            }
        }
    }
    // output unit: dot(w4, h3) over the 64 units (4 lanes of a leaf hold 16 each)
    float part = 0.f;
#pragma unroll
    for (int mt = 0; mt < 4; mt++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float h = gelu(o.z[2][mt][j]);
            o.h3[mt][j] = h;
            part += sB[SB_W4 + 16 * mt + 4 * g + j] * h;
        }
    part = sum4_rows(part);
    o.z4 = part + sB[SB_B4];
}

__device__ __forceinline__ void load_weights(const uint16_t* __restrict__ wpk, const float* __restrict__ w32,
                                             uint16_t* sW, float* sB, int n_bf) {
    for (int k = threadIdx.x; k < n_bf; k += blockDim.x) sW[k] = wpk[k];
    for (int k = threadIdx.x; k < MLP_H; k += blockDim.x) {
        sB[SB_B1 + k] = w32[P_B1 + k];
        sB[SB_B2 + k] = w32[P_B2 + k];
        sB[SB_B3 + k] = w32[P_B3 + k];
        sB[SB_W4 + k] = w32[P_W4 + k];
    }
    if (threadIdx.x == 0) sB[SB_B4] = w32[P_B4];
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352dU; x ^= x >> 15; x *= 0x846ca68bU; x ^= x >> 16;
    return x;
}

// ---------------------------------------------------------------- inference + judge sampling
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 frag(const uint16_t* sF, int f) {
    return *reinterpret_cast<const bf16x8*>(sF + (f * 64 + (threadIdx.x & 63)) * 8);
}
// GELU of two 16-unit accumulator tiles packed as the B fragment of the next layer: slots
// 0-3 hold units 32s + 4g + j, slots 4-7 units 32s + 16 + 4g + j (the weight fragments use
// the same K order).
__device__ __forceinline__ bf16x8 act_pack(const f32x4& lo, const f32x4& hi) {
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        r[j] = (__bf16)gelu(lo[j]);
        r[4 + j] = (__bf16)gelu(hi[j]);
    }
    return r;
}

// flags bit 0: troubled (full path); bit 1: judged this step (cleared by the network but
// sent through the full path with a label); bit 2: forced troubled by an earlier judge
// rejection (sticky until the tables are rebuilt); bit 3: made by the last regrid (refined
// son or coarsened parent: copied or averaged values) or reading such a leaf in its stencil,
// troubled in the first step after the regrid. Bits 0 and 1 hold for the whole step (the
// judge only adds bit 2), so the stage launches can read the troubled set from the flags.
__device__ __forceinline__ uint8_t decide(uint8_t old, bool troubled, int leaf, const SelectParams& sp) {
    uint8_t f = old & 4;
    if (troubled || f || (old & 8)) f |= 1;
    else if ((hash32((uint32_t)leaf * 2654435761u ^ sp.step) & 0xFFFFu) < sp.judge_q16) f |= 3;
    return f;
}

// Inference in one pass: a persistent grid whose blocks hold the network's weights in LDS (as
// lane-contiguous fragments) and whose waves each take one 16-leaf tile at a time: the channels
// of the tile's distinct neighbourhood states go to the wave's LDS region (each converted
// once), the 5x5 features come from there, near-uniform leaves are cleared without the
// network, and for the others the gathered features are already the layer-1 B fragments
// (feature q = 8s + j of lane (r, g) is channel q % 3 of point g + 4 (q / 3)), so the network
// runs on them in registers: every layer's accumulators are the next layer's B fragments. Each XCD takes a contiguous eighth of the
// tiles (its waves stride over it), so neighbouring tiles share that XCD's L2.
// Waves per block: a multiple of 4, so the two blocks on a CU put the same number of waves
// on each SIMD (10-wave blocks left two SIMDs with 6 waves and two with 4, and ran 1.5x
// longer); the LDS of a block (weights + 12 neighbourhoods) is 66.6 KB.
constexpr int SF_WAVES = 12;
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Neighbourhood extent of a tile (wave-uniform: scalar loads) and what the pipeline of
// k_select fetches ahead for a tile: the lanes' position words, the flag byte, the indices of
// neighbourhood entries lane and lane + 64; then their FP32 states, straight into the wave's
// LDS region (global_load_lds: lane l's 16 bytes land at entry l of the region).
struct TileMeta { int b, n; };
__device__ __forceinline__ TileMeta tile_meta(const int32_t* __restrict__ nbr_beg, int tile) {
    const int b = nbr_beg[tile];
    return TileMeta{b, nbr_beg[tile + 1] - b};
}
struct TilePre {
    uint2 lc;
    uint32_t old;   // the leaf's flag byte (every lane loads it; lanes g > 0 ignore it)
    int idx[2];
};
// Every load is unconditional (a lane past the extent reads an in-range entry it then ignores),
// so no branch makes the wave wait for the data before it is used. nbr_end: the length of nbr.
__device__ __forceinline__ void tile_pre_index(const DevTables& t, const int32_t* __restrict__ nbr, int nbr_end,
                                               const uint2* __restrict__ loc, const uint8_t* flags, int tile,
                                               const TileMeta& m, TilePre& p) {
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int leaf = tile * 16 + r;
    const int lc = leaf < t.N ? leaf : 0;
    // 32-bit unsigned offsets from the (scalar) bases: no per-lane 64-bit addresses to keep
    p.lc = loc[(uint32_t)(lc * 4 + g)];
    p.old = flags[(uint32_t)lc];
    p.idx[0] = nbr[(uint32_t)min(m.b + lane, nbr_end - 1)];
    p.idx[1] = nbr[(uint32_t)min(m.b + lane + 64, nbr_end - 1)];
}
typedef __attribute__((address_space(3))) void lds_void;
__device__ __forceinline__ void tile_pre_states(const float4* __restrict__ Qf, const TilePre& p, float4* region) {
    __builtin_amdgcn_global_load_lds((const void*)(Qf + p.idx[0]), (lds_void*)region, 16, 0, 0);
    __builtin_amdgcn_global_load_lds((const void*)(Qf + p.idx[1]), (lds_void*)(region + 64), 16, 0, 0);
}
// one hidden layer of the network on a 16-leaf tile: z <- W^T gelu(z) + b (K = 64)
__device__ __forceinline__ void net_layer(const uint16_t* W, const float* bias, f32x4 z[4]) {
    const int g = (threadIdx.x & 63) >> 4;
    const bf16x8 hb0 = act_pack(z[0], z[1]), hb1 = act_pack(z[2], z[3]);
#pragma unroll
    for (int mt = 0; mt < 4; mt++) {
        f32x4 acc = *reinterpret_cast<const f32x4*>(bias + 16 * mt + 4 * g);
        acc = mfma32(frag(W, mt * 2 + 0), hb0, acc);
        acc = mfma32(frag(W, mt * 2 + 1), hb1, acc);
        z[mt] = acc;
    }
}

__global__ __launch_bounds__(SF_WAVES * 64) __attribute__((amdgpu_waves_per_eu(6))) void k_select(DevTables t, const float4* __restrict__ Qf,
                                                          const int32_t* __restrict__ off16, SelectParams sp,
                                                          uint8_t* __restrict__ flags, float* __restrict__ prob,
                                                          const int32_t* __restrict__ nbr_beg,
                                                          const int32_t* __restrict__ nbr,
                                                          const uint2* __restrict__ loc,
                                                          const uint16_t* __restrict__ wpk,
                                                          const float* __restrict__ w32,
                                                          int* __restrict__ counters) {
    // one LDS array with the weight fragments at its start, so every fragment read encodes
    // its offset in the instruction (the offset field reaches 64 KB; the neighbourhoods come last)
    constexpr int SB_BYTES = (SB_TOTAL * 4 + 15) / 16 * 16;
    __shared__ __align__(16) unsigned char lds[KF_TOTAL * 2 + SB_BYTES + SF_WAVES * SEL_NBR_CAP * 16];
    uint16_t* sF = reinterpret_cast<uint16_t*>(lds);
    float* sB = reinterpret_cast<float*>(lds + KF_TOTAL * 2);
    float4(*sQ)[SEL_NBR_CAP] = reinterpret_cast<float4(*)[SEL_NBR_CAP]>(lds + KF_TOTAL * 2 + SB_BYTES);
    {
        const uint4* src = reinterpret_cast<const uint4*>(wpk + KW_FRAG);
        uint4* dst = reinterpret_cast<uint4*>(sF);
        for (int k = threadIdx.x; k < KF_TOTAL / 8; k += blockDim.x) dst[k] = src[k];
        for (int k = threadIdx.x; k < MLP_H; k += blockDim.x) {
            sB[SB_B1 + k] = w32[P_B1 + k];
            sB[SB_B2 + k] = w32[P_B2 + k];
            sB[SB_B3 + k] = w32[P_B3 + k];
            sB[SB_W4 + k] = w32[P_W4 + k];
        }
        if (threadIdx.x == 0) sB[SB_B4] = w32[P_B4];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int ntiles = (t.N + 15) / 16;
    // this block's XCD range of tiles (the grid is a multiple of 8 blocks)
    const int x = blockIdx.x & 7, xb = blockIdx.x >> 3, nbx = gridDim.x >> 3;
    const int lo = (int)(((long long)ntiles * x) >> 3), hi = (int)(((long long)ntiles * (x + 1)) >> 3);
    const int stride = nbx * SF_WAVES;
    int net_tiles = 0;  // tiles this wave ran the network on
    // Software pipeline over the wave's tiles T, T + stride, ...: while tile T's features and
    // network run, the next tile's position words, old flags and neighbourhood indices are
    // loaded (before layer 1), then its first 128 neighbourhood states (before layer 3), and
    // the neighbourhood extent of the tile after it; so the dependent loads of a tile
    // (extent -> indices -> states) overlap the network of the one before.
    int tile = lo + xb * SF_WAVES + wv;
    TileMeta mc = tile < hi ? tile_meta(nbr_beg, tile) : TileMeta{0, 0};
    TileMeta mn = tile + stride < hi ? tile_meta(nbr_beg, tile + stride) : TileMeta{0, 0};
    // entries of the marker tiles (gathered directly) are 0; every index is a state index
    const int nbr_end = nbr_beg[ntiles];
    TilePre pc;
    if (tile < hi) {
        tile_pre_index(t, nbr, nbr_end, loc, flags, tile, mc, pc);
        tile_pre_states(Qf, pc, sQ[wv]);
    }
    for (; tile < hi; tile += stride) {
        const int leaf = tile * 16 + r;
        const bool valid = leaf < t.N;
        const bool local = mc.n <= SEL_NBR_CAP;
        const uint2 lc = pc.lc;
        uint32_t old32 = valid && g == 0 ? pc.old & 0xFFu : 0u;
        // materialised here, where the wave waits for this tile's data anyway: used only after
        // the next tile's loads are issued, it would make the wave wait for those too
        asm volatile("" : "+v"(old32));
        const uint8_t old = (uint8_t)old32;
        if (local) {
            // entries 0..127 arrived as raw states: converted in place (each lane its own two)
            if (lane < mc.n) sQ[wv][lane] = channels(sQ[wv][lane]);
            if (lane + 64 < mc.n) sQ[wv][lane + 64] = channels(sQ[wv][lane + 64]);
            for (int k = lane + 128; k < mc.n; k += 64) sQ[wv][k] = channels(Qf[nbr[mc.b + k]]);
        }
        wave_lds_sync();
        float v[7][3];
        bool uniform;
        if (local) {
            // positions are below nU (checked when the tables are built); leaves past N use entry 0
            const uint8_t* pos = reinterpret_cast<const uint8_t*>(&lc);
            float4 q[7];
#pragma unroll
            for (int k = 0; k < 7; k++) q[k] = sQ[wv][valid ? (int)pos[k] : 0];
            // byte 7: the centre's position
            const float4 qc = sQ[wv][valid ? (int)pos[7] : 0];
            uniform = features_from(qc, q, v);
        } else {
            int idx[7];
            gather_idx(t, off16, leaf, valid, idx, true);
            uniform = gather_values(Qf, leaf, valid, idx, v);
        }
        // the next tile's neighbourhood overwrites this one only after every lane read it
        wave_lds_sync();
        // next tile: position words, old flags, neighbourhood indices; the extent of the one after
        const int tn = tile + stride;
        TilePre pn;
        if (tn < hi) tile_pre_index(t, nbr, nbr_end, loc, flags, tn, mn, pn);
        const TileMeta mnn = tn + stride < hi ? tile_meta(nbr_beg, tn + stride) : TileMeta{0, 0};
        const uint64_t need = __ballot(valid && !uniform);
        const uint16_t mask = (uint16_t)(need & 0xFFFFu);
        if (valid && g == 0 && uniform) {
            if (prob) prob[leaf] = 0.f;
            flags[leaf] = decide(old, false, leaf, sp);
        }
        f32x4 z[4];
        if (mask) {
            net_tiles++;
            bf16x8 xb3[3];
#pragma unroll
            for (int s3 = 0; s3 < 3; s3++) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int q = 8 * s3 + j;
                    xb3[s3][j] = (__bf16)(q < 21 ? v[q / 3][q % 3] : 0.f);
                }
            }
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                z[mt] = *reinterpret_cast<const f32x4*>(sB + SB_B1 + 16 * mt + 4 * g);
#pragma unroll
                for (int s3 = 0; s3 < 3; s3++) z[mt] = mfma32(frag(sF + KF_L1, mt * 3 + s3), xb3[s3], z[mt]);
            }
            net_layer(sF + KF_L2, sB + SB_B2, z);
        }
        // next tile: its first 128 neighbourhood states
        if (tn < hi) tile_pre_states(Qf, pn, sQ[wv]);
        if (mask) {
            net_layer(sF + KF_L3, sB + SB_B3, z);
            float part = 0.f;
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                const f32x4 w4 = *reinterpret_cast<const f32x4*>(sB + SB_W4 + 16 * mt + 4 * g);
#pragma unroll
                for (int j = 0; j < 4; j++) part += w4[j] * gelu(z[mt][j]);
            }
            part = sum4_rows(part);
            if (valid && g == 0 && ((mask >> r) & 1)) {
                const float pr = 1.f / (1.f + __expf(-(part + sB[SB_B4])));
                if (prob) prob[leaf] = pr;
                flags[leaf] = decide(old, pr >= sp.threshold, leaf, sp);
            }
        }
        mc = mn;
        mn = mnn;
        pc = pn;
    }
    if (counters && lane == 0 && net_tiles)
        atomicAdd(&counters[(blockIdx.x % CNT_SLOTS) * CNT_N + CNT_TILES], net_tiles);
}

// judge: a judged leaf whose label exceeds the bound is forced back to the full path; also
// counts the leaves that took the full path (bit 0). Each block scans 1024 leaves (4 per
// thread) and adds its counts with one atomic per counter.
__global__ __launch_bounds__(256) void k_judge(int n, const float* __restrict__ label, float bound,
                                               uint8_t* __restrict__ flags, int* __restrict__ counters) {
    int judged = 0, rejected = 0, troubled = 0;
    const int base = blockIdx.x * 1024;
    for (int k = threadIdx.x; k < 1024; k += 256) {
        const int i = base + k;
        if (i >= n) break;
        const uint8_t f = flags[i];
        troubled += f & 1;
        if (f & 2) {
            judged++;
            // bits 0 and 1 stay (the leaf is on the full path for the whole step)
            if (label[i] > bound) { rejected++; flags[i] = (uint8_t)(f | 4); }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        judged += __shfl_xor(judged, o);
        rejected += __shfl_xor(rejected, o);
        troubled += __shfl_xor(troubled, o);
    }
    __shared__ int sj[4], sr[4], st[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sj[wv] = judged; sr[wv] = rejected; st[wv] = troubled; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int j = sj[0] + sj[1] + sj[2] + sj[3], r = sr[0] + sr[1] + sr[2] + sr[3];
        const int tr = st[0] + st[1] + st[2] + st[3];
        int* slot = counters + (blockIdx.x % CNT_SLOTS) * CNT_N;
        if (j) atomicAdd(&slot[CNT_JUDGED], j);
        if (r) atomicAdd(&slot[CNT_REJECTED], r);
        if (tr) atomicAdd(&slot[CNT_TROUBLED], tr);
    }
}

// features of every leaf to global (training data / tests)
__global__ __launch_bounds__(256) void k_features(DevTables t, const float4* __restrict__ Qf,
                                                  const int32_t* __restrict__ off16, uint16_t* __restrict__ X,
                                                  int row0) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int tile = blockIdx.x * 4 + wv;
    const int leaf = tile * 16 + (lane & 15);
    const bool valid = leaf < t.N;
    __shared__ uint16_t sX[4][16 * MLP_IN];
    leaf_features(t, Qf, off16, leaf, valid, sX[wv] + (lane & 15) * MLP_IN);
    __syncthreads();
    // lanes copy the tile out (16 rows of 80)
    for (int k = lane; k < 16 * MLP_IN; k += 64) {
        const int rr = k / MLP_IN;
        const int lf = tile * 16 + rr;
        if (lf < t.N) X[(size_t)(row0 + lf) * MLP_IN + (k % MLP_IN)] = sX[wv][k];
    }
}

__global__ void k_labels_to_y(int n, const float* __restrict__ label, float bound, uint8_t* __restrict__ y,
                              int row0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[row0 + i] = label[i] > bound ? 1 : 0;
}

// ---------------------------------------------------------------- training
// One wave per block; the block accumulates its gradient privately in gpart[blockIdx.x] (no
// atomics: every gradient element of a tile is owned by one lane), reduced afterwards.
__global__ __launch_bounds__(64) void k_mlp_grad(const uint16_t* __restrict__ wpk, const float* __restrict__ w32,
                                                 const uint16_t* __restrict__ X, const uint8_t* __restrict__ y,
                                                 const int32_t* __restrict__ rows, int n, float pos_weight,
                                                 float scale, float* __restrict__ gpart, float* __restrict__ lpart) {
    __shared__ uint16_t sW[KW_PLAIN];
    __shared__ float sB[SB_TOTAL];
    __shared__ uint16_t sX[16 * MLP_IN];
    __shared__ uint16_t sH1[MLP_H * 16], sH2[MLP_H * 16], sDZ[MLP_H * 16];
    load_weights(wpk, w32, sW, sB, KW_PLAIN);
    float* gp = gpart + (size_t)blockIdx.x * P_TOTAL;
    for (int k = threadIdx.x; k < P_TOTAL; k += 64) gp[k] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const int ntiles = (n + 15) / 16;
    float loss = 0.f;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int idx = tile * 16 + r;
        const bool valid = idx < n;
        const int row = valid ? (rows ? rows[idx] : idx) : 0;
        for (int k = g; k < MLP_IN / 4; k += 4) {
            const uint16_t* src = X + (size_t)row * MLP_IN + 4 * k;
#pragma unroll
            for (int j = 0; j < 4; j++) sX[r * MLP_IN + 4 * k + j] = valid ? src[j] : 0;
        }
        __syncthreads();
        FwdOut o;
        mlp_forward(sW, sB, sX, o, sH1, sH2);
        const float yv = valid ? (float)y[row] : 0.f;
        const float p = 1.f / (1.f + expf(-o.z4));
        const float pc = fminf(fmaxf(p, 1e-7f), 1.f - 1e-7f);
        if (valid && g == 0) loss += -(pos_weight * yv * logf(pc) + (1.f - yv) * logf(1.f - pc));
        const float dz4 = valid ? scale * (p * (pos_weight * yv + 1.f - yv) - pos_weight * yv) : 0.f;
        // output layer gradients
        {
            float gb = g == 0 ? dz4 : 0.f;
            for (int o2 = 8; o2 > 0; o2 >>= 1) gb += __shfl_xor(gb, o2);
            if (lane == 0) gp[P_B4] += gb;
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    float v = o.h3[mt][j] * dz4;
                    for (int o2 = 8; o2 > 0; o2 >>= 1) v += __shfl_xor(v, o2);
                    if (r == 0) gp[P_W4 + 16 * mt + 4 * g + j] += v;
                }
        }
        // dz3 = w4 dz4 gelu'(z3)
        f32x4 dz[4];
#pragma unroll
        for (int mt = 0; mt < 4; mt++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                dz[mt][j] = sB[SB_W4 + 16 * mt + 4 * g + j] * dz4 * gelu_d(o.z[2][mt][j]);
        // layers 3, 2, 1 (L = 2, 1, 0)
#pragma unroll
        for (int L = 2; L >= 0; L--) {
            const int pw = L == 2 ? P_W3 : L == 1 ? P_W2 : P_W1;
            const int pb = L == 2 ? P_B3 : L == 1 ? P_B2 : P_B1;
            // bias gradient: sum over the tile's leaves
#pragma unroll
            for (int mt = 0; mt < 4; mt++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    float v = dz[mt][j];
                    for (int o2 = 8; o2 > 0; o2 >>= 1) v += __shfl_xor(v, o2);
                    if (r == 0) gp[pb + 16 * mt + 4 * g + j] += v;
                }
            // stage dz (BF16) in [unit][leaf] layout
            bf16x4 dzb[4];
#pragma unroll
            for (int mt = 0; mt < 4; mt++) {
                dzb[mt] = pack4(dz[mt][0], dz[mt][1], dz[mt][2], dz[mt][3]);
#pragma unroll
                for (int j = 0; j < 4; j++) sDZ[(16 * mt + 4 * g + j) * 16 + r] = (uint16_t)dzb[mt][j];
            }
            __syncthreads();
            // weight gradient dW[in][out] = sum_leaf act[in][leaf] dz[out][leaf] (K = 16 leaves)
            const int nin_t = L == 0 ? 5 : 4;
            for (int mi = 0; mi < nin_t; mi++) {
                bf16x4 a;
                if (L == 0) {
#pragma unroll
                    for (int j = 0; j < 4; j++) a[j] = (short)sX[(4 * g + j) * MLP_IN + 16 * mi + r];
                } else {
                    const uint16_t* act = L == 2 ? sH2 : sH1;
                    a = lds4(act + (16 * mi + r) * 16 + 4 * g);
                }
#pragma unroll
                for (int no = 0; no < 4; no++) {
                    const bf16x4 b = lds4(sDZ + (16 * no + r) * 16 + 4 * g);
                    f32x4 c = {0.f, 0.f, 0.f, 0.f};
                    c = mfma16(a, b, c);
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int in = 16 * mi + 4 * g + j, out = 16 * no + r;
                        if (in < 75) gp[pw + in * MLP_H + out] += c[j];
                    }
                }
            }
            if (L > 0) {
                // dh_{L-1}[in][leaf] = W_L[in][out] dz[out][leaf]; then dz_{L-1} = dh gelu'(z_{L-1})
                const int wn = L == 2 ? KW_W3N : KW_W2N;
                f32x4 nd[4];
#pragma unroll
                for (int mt = 0; mt < 4; mt++) {
                    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        const bf16x4 a = lds4(sW + wn + (16 * mt + r) * MLP_H + 16 * s + 4 * g);
                        acc = mfma16(a, dzb[s], acc);
                    }
                    nd[mt] = acc;
                }
#pragma unroll
                for (int mt = 0; mt < 4; mt++)
#pragma unroll
                    for (int j = 0; j < 4; j++) dz[mt][j] = nd[mt][j] * gelu_d(o.z[L - 1][mt][j]);
            }
            __syncthreads();
        }
    }
    for (int o2 = 32; o2 > 0; o2 >>= 1) loss += __shfl_xor(loss, o2);
    if (lane == 0) lpart[blockIdx.x] = loss;
}

__global__ void k_reduce_grad(const float* __restrict__ gpart, int nb, const float* __restrict__ lpart,
                              float* __restrict__ grad, float* __restrict__ loss) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < P_TOTAL) {
        float s = 0.f;
        for (int b = 0; b < nb; b++) s += gpart[(size_t)b * P_TOTAL + k];
        grad[k] = s;
    }
    if (k == 0) {
        float s = 0.f;
        for (int b = 0; b < nb; b++) s += lpart[b];
        *loss = s;
    }
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float bc1, float bc2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= P_TOTAL) return;
    if (k < P_B1 && (k / MLP_H) >= 75) return;  // padded input rows stay zero
    const float gk = grad[k];
    const float mk = b1 * m[k] + (1.f - b1) * gk;
    const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    w[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
}

// BF16 copies for the MFMA operands: W1T [64][80], W2T, W3T [64][64] (transposed: A operand of
// the forward pass) and W2N, W3N [64][64] (as stored: A operand of the backward data path)
__global__ void k_pack(const float* __restrict__ w, uint16_t* __restrict__ pk) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < MLP_H * MLP_IN) {
        const int o = k / MLP_IN, i = k % MLP_IN;
        pk[KW_W1T + k] = f2bf(w[P_W1 + i * MLP_H + o]);
    }
    if (k < MLP_H * MLP_H) {
        const int o = k / MLP_H, i = k % MLP_H;
        pk[KW_W2T + k] = f2bf(w[P_W2 + i * MLP_H + o]);
        pk[KW_W3T + k] = f2bf(w[P_W3 + i * MLP_H + o]);
        pk[KW_W2N + k] = f2bf(w[P_W2 + k]);
        pk[KW_W3N + k] = f2bf(w[P_W3 + k]);
    }
    if (k < KF_TOTAL) {
        // fragment order of the inference kernel: element j of lane (r, g) in fragment f
        const int j = k & 7, lane = (k >> 3) & 63, r = lane & 15, g = lane >> 4;
        float val = 0.f;
        if (k < KF_L2) {
            const int f = k / KF_FRAG, mt = f / 3, q = (f % 3) * 8 + j, p = g + 4 * (q / 3);
            if (q < 21 && p < 25) val = w[P_W1 + (p * 3 + q % 3) * MLP_H + 16 * mt + r];
        } else {
            const int base = k < KF_L3 ? P_W2 : P_W3;
            const int f = ((k < KF_L3 ? k - KF_L2 : k - KF_L3)) / KF_FRAG, mt = f / 2, s2 = f % 2;
            const int unit = 32 * s2 + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4);
            val = w[base + unit * MLP_H + 16 * mt + r];
        }
        pk[KW_FRAG + k] = f2bf(val);
    }
}

// ---------------------------------------------------------------- launchers
void launch_select(const DevTables& t, const float4* Qf, const int32_t* off16, const uint16_t* wpk,
                   const float* w32, const SelectParams& sp, uint8_t* flags, float* prob, int n_cu,
                   const int32_t* nbr_beg, const int32_t* nbr, const uint8_t* loc, int* counters,
                   hipStream_t st) {
    if (t.N <= 0) return;
    // two blocks per CU (the LDS limit), a multiple of 8 (one tile range per XCD)
    const int blocks = ((2 * n_cu + 7) / 8) * 8;
    hipLaunchKernelGGL(k_select, dim3(blocks), dim3(SF_WAVES * 64), 0, st, t, Qf, off16, sp, flags, prob, nbr_beg,
                       nbr, reinterpret_cast<const uint2*>(loc), wpk, w32, counters);
}

void launch_judge(int n, const float* label, float bound, uint8_t* flags, int* counters, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_judge, dim3((n + 1023) / 1024), dim3(256), 0, st, n, label, bound, flags, counters);
}

void launch_features(const DevTables& t, const float4* Qf, const int32_t* off16, uint16_t* X, int row0,
                     hipStream_t st) {
    if (t.N <= 0) return;
    const int ntiles = (t.N + 15) / 16;
    hipLaunchKernelGGL(k_features, dim3((ntiles + 3) / 4), dim3(256), 0, st, t, Qf, off16, X, row0);
}

void launch_labels_to_y(int n, const float* label, float bound, uint8_t* y, int row0, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_labels_to_y, dim3((n + 255) / 256), dim3(256), 0, st, n, label, bound, y, row0);
}

void launch_mlp_grad(const uint16_t* wpk, const float* w32, const uint16_t* X, const uint8_t* y,
                     const int32_t* rows, int n, float pos_weight, float scale, float* gpart, float* lpart,
                     int nblocks, float* grad, float* loss, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(64), 0, st, wpk, w32, X, y, rows, n, pos_weight,
                       scale, gpart, lpart);
    hipLaunchKernelGGL(k_reduce_grad, dim3((P_TOTAL + 255) / 256), dim3(256), 0, st, gpart, nblocks, lpart,
                       grad, loss);
}

void launch_adam(float* w, const float* grad, float* m, float* v, float lr, float b1, float b2, float eps,
                 float bc1, float bc2, hipStream_t st) {
    hipLaunchKernelGGL(k_adam, dim3((P_TOTAL + 255) / 256), dim3(256), 0, st, w, grad, m, v, lr, b1, b2, eps,
                       bc1, bc2);
}

void launch_pack(const float* w, uint16_t* pk, hipStream_t st) {
    hipLaunchKernelGGL(k_pack, dim3((KF_TOTAL + 255) / 256), dim3(256), 0, st, w, pk);
}

}  // namespace fdm
