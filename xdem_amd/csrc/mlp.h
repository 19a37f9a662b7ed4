// mlp.h -- the learned quartic-root guess: an 8-64-64-1 ReLU network on the
// gfx950 matrix cores (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulate), shared by the step kernels and the training kernels.
//
// No counterpart in the reference: its Newton iteration always starts at
// sqrt(delta + 1) (src/EllipticParticle.cpp:163).  The network maps eight
// pose/shape features of an ellipse pair to u = log q / log sqrt(delta + 1),
// the place of the root q between 1 and the plain start value.
//
// One wavefront evaluates a tile of 16 pairs.  The layers are computed
// transposed, H^T = W X^T, so that the hidden units are the MFMA's M
// dimension and the pairs its N dimension.  Lane l then holds, for pair
// l % 16, the hidden units 16 t + 4 (l / 16) + e (t, e = 0..3) in its four
// accumulator tiles, and exactly those values form its B operand of the next
// layer once the next layer's weights are stored with the matching order of
// their K index.  Activations never leave the lane's registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xdem {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kMlpIn = 8;
constexpr int kMlpHid = 64;
// FP32 parameter vector: W1[64][8], b1[64], W2[64][64], b2[64], w3[64], b3
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kMlpHid * kMlpIn;
constexpr int kOffW2 = kOffB1 + kMlpHid;
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid;
constexpr int kOffW3 = kOffB2 + kMlpHid;
constexpr int kOffB3 = kOffW3 + kMlpHid;
constexpr int kMlpParams = kOffB3 + 1;       // 4801

// MFMA 16x16x32 operand layout: lane l holds A[l % 16][8 (l / 16) + s] and
// B[8 (l / 16) + s][l % 16] for s = 0..7, and D[4 (l / 16) + e][l % 16] for
// e = 0..3.  Hidden unit of K slot s of chunk c in lane group g (layers 2
// and the backward product): the accumulator value (t = 2c + s/4, e = s%4).
__host__ __device__ constexpr int mlp_k_unit(int c, int g, int s) {
    return 16 * (2 * c + s / 4) + 4 * g + (s % 4);
}

// Weights as per-lane MFMA operands, refreshed from the FP32 parameters
// after every optimizer step.
struct MlpPacked {
    bf16x8 a1[4][64];        // layer 1, tile t: W1[16t + l%16][f = 8 (l/16) + s] (f < 8)
    bf16x8 a2[4][2][64];     // layer 2, tile m, chunk c: W2[16m + l%16][unit(c, l/16, s)]
    bf16x8 a2t[4][2][64];    // backward: W2[unit(c, l/16, s)][16m + l%16]
    float b1[kMlpHid], b2[kMlpHid], w3[kMlpHid], b3;
};

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// The lane's share of one network evaluation: the values it holds of the
// 16-pair tile.  h1[4t + e], h2[4m + e] are hidden units 16t + 4g + e.
struct MlpLane {
    float h1[16];
    float h2[16];
    float y;                 // network output of pair l % 16 (every lane)
};

// Forward pass of a 16-pair tile.  x: the lane's B operand of layer 1, the
// pair's 8 features in lanes 0-15 and zeros elsewhere.  `W` holds the lane's
// operands (already loaded), `bias1/bias2/w3v` its 16 units' values.
__device__ __forceinline__ void mlp_forward(const bf16x8 x, const bf16x8 (&a1)[4],
                                            const bf16x8 (&a2)[4][2], const float (&bias1)[16],
                                            const float (&bias2)[16], const float (&w3v)[16],
                                            float b3, MlpLane& L) {
    const f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const f32x4 d = mfma16(a1[t], x, z);
#pragma unroll
        for (int e = 0; e < 4; ++e) L.h1[4 * t + e] = fmaxf(d[e] + bias1[4 * t + e], 0.f);
    }
    bf16x8 bq[2];
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int s = 0; s < 8; ++s) bq[c][s] = (__bf16)L.h1[8 * c + s];
    float part = 0.f;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        f32x4 d = mfma16(a2[m][0], bq[0], z);
        d = mfma16(a2[m][1], bq[1], d);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float h = fmaxf(d[e] + bias2[4 * m + e], 0.f);
            L.h2[4 * m + e] = h;
            part += w3v[4 * m + e] * h;
        }
    }
    part += __shfl_xor(part, 16, 64);
    part += __shfl_xor(part, 32, 64);
    L.y = part + b3;
}

// Load the lane's operands and bias values from the packed weights.
struct MlpRegs {
    bf16x8 a1[4];
    bf16x8 a2[4][2];
    float b1[16], b2[16], w3[16];
    float b3;
};

__device__ __forceinline__ void mlp_load(const MlpPacked* __restrict__ W, MlpRegs& R) {
    const int l = threadIdx.x & 63, g = l >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) R.a1[t] = W->a1[t][l];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 2; ++c) R.a2[m][c] = W->a2[m][c][l];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int u = 16 * t + 4 * g + e;
            R.b1[4 * t + e] = W->b1[u];
            R.b2[4 * t + e] = W->b2[u];
            R.w3[4 * t + e] = W->w3[u];
        }
    R.b3 = W->b3;
}

// Features of an ellipse pair for the root guess, all O(1): the direction of
// the centre line d and the orientation of P2, in P1's frame, as doubled
// angles (each is defined mod pi); the direction of d in P2's frame; the
// shape ratios b1/a1, a2/a1, b2/a1.  k1, k2: the major-axis directions.
__device__ __forceinline__ void root_features(float dx, float dy, float k1x, float k1y,
                                              float k2x, float k2y, float a1, float b1,
                                              float a2, float b2, float (&f)[kMlpIn]) {
    const float il = 1.0f / sqrtf(dx * dx + dy * dy);
    const float ca = (dx * k1x + dy * k1y) * il, sa = (dy * k1x - dx * k1y) * il;
    const float cb = k1x * k2x + k1y * k2y, sb = k1x * k2y - k1y * k2x;
    const float c2a = ca * ca - sa * sa, s2a = 2.0f * ca * sa;
    const float c2b = cb * cb - sb * sb, s2b = 2.0f * cb * sb;
    f[0] = c2a;
    f[1] = s2a;
    f[2] = c2b;
    f[3] = s2b;
    f[4] = c2a * c2b + s2a * s2b;
    f[5] = b1 / a1;
    f[6] = a2 / a1;
    f[7] = b2 / a1;
}

__device__ __forceinline__ uint4 pack_features(const float (&f)[kMlpIn]) {
    bf16x8 v;
#pragma unroll
    for (int s = 0; s < 8; ++s) v[s] = (__bf16)f[s];
    return *reinterpret_cast<const uint4*>(&v);
}

__device__ __forceinline__ bf16x8 unpack_features(uint4 u) {
    return *reinterpret_cast<const bf16x8*>(&u);
}

}  // namespace xdem
