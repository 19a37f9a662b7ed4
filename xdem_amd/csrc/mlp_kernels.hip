// mlp_kernels.hip -- training and inference kernels of the learned
// quartic-root guess (the network itself is in mlp.h).
//
//   k_mlp_pack    FP32 parameters -> per-lane BF16 MFMA operands
//   k_mlp_rows    inference over a row array (checks, hold-out evaluation)
//   k_mlp_train   forward + backward over a mini-batch; one partial gradient
//                 (and squared-error sum) per wavefront
//   k_mlp_adam    fixed-order sum of the partials; Adam step (or only the
//                 gradient, for checks)
//
// Loss: mean squared error of u over the batch.  Rounding points (mirrored
// by the CPU oracle of the network): features, W1, W2 and the layer-1
// activations are rounded to BF16 for the MFMA products; the layer-2 delta
// is rounded to BF16 for the backward product through W2; sums are FP32.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp.h"
#include "mlp_launch.h"

namespace xdem {

namespace {

constexpr int kWaves = 4;                 // wavefronts per training block
constexpr int kPartial = kMlpParams + 1;  // gradient + squared-error sum

__global__ void k_mlp_pack(const float* __restrict__ prm, MlpPacked* __restrict__ pk) {
    const int l = threadIdx.x;            // 64 lanes
    const int g = l >> 4, r = l & 15;
    for (int t = 0; t < 4; ++t) {
        bf16x8 v;
        for (int s = 0; s < 8; ++s) {
            const int f = 8 * g + s;
            v[s] = (__bf16)(f < kMlpIn ? prm[kOffW1 + (16 * t + r) * kMlpIn + f] : 0.0f);
        }
        pk->a1[t][l] = v;
    }
    for (int m = 0; m < 4; ++m)
        for (int c = 0; c < 2; ++c) {
            bf16x8 v, w;
            for (int s = 0; s < 8; ++s) {
                const int u = mlp_k_unit(c, g, s);
                v[s] = (__bf16)prm[kOffW2 + (16 * m + r) * kMlpHid + u];
                w[s] = (__bf16)prm[kOffW2 + u * kMlpHid + 16 * m + r];
            }
            pk->a2[m][c][l] = v;
            pk->a2t[m][c][l] = w;
        }
    pk->b1[l] = prm[kOffB1 + l];
    pk->b2[l] = prm[kOffB2 + l];
    pk->w3[l] = prm[kOffW3 + l];
    if (l == 0) pk->b3 = prm[kOffB3];
}

// Inference over rows: one wavefront per 16-row tile (grid-stride).
__global__ void __launch_bounds__(256)
k_mlp_rows(const uint4* __restrict__ feat, int rows, const MlpPacked* __restrict__ pk,
           float* __restrict__ out) {
    const int l = threadIdx.x & 63, g = l >> 4, r = l & 15;
    MlpRegs R;
    mlp_load(pk, R);
    const int ntiles = (rows + 15) / 16;
    const int nw = gridDim.x * (blockDim.x >> 6);
    for (int tile = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); tile < ntiles; tile += nw) {
        const int row = tile * 16 + r;
        bf16x8 x = {};
        if (g == 0 && row < rows) x = unpack_features(feat[row]);
        MlpLane L;
        mlp_forward(x, R.a1, R.a2, R.b1, R.b2, R.w3, R.b3, L);
        if (g == 0 && row < rows) out[row] = L.y;
    }
}

// Forward and backward over the rows idx[0..rows) (idx == nullptr: 0..rows).
// dscale = 2 / (batch size): d loss / d y of a row is dscale * (y - target).
// The waves of a block run the same number of rounds (tiles of 16 rows, a
// tile past the end holds no rows) so that the block barriers are uniform.
__global__ void __launch_bounds__(256)
k_mlp_train(const uint4* __restrict__ feat, const float* __restrict__ targ,
            const int* __restrict__ idx, int rows, float dscale,
            const MlpPacked* __restrict__ pk, float* __restrict__ partial) {
    __shared__ float s_h1[kWaves][16][kMlpHid];     // BF16-rounded layer-1 activations
    __shared__ float s_h2[kWaves][16][kMlpHid];
    __shared__ float s_d2[kWaves][16][kMlpHid];     // d loss / d z2
    __shared__ float s_d1[kWaves][16][kMlpHid];     // d loss / d z1
    __shared__ float s_x[kWaves][16][kMlpIn];
    __shared__ float s_dy[kWaves][16];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63, g = l >> 4, r = l & 15;
    MlpRegs R;
    mlp_load(pk, R);
    bf16x8 a2t[4][2];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int c = 0; c < 2; ++c) a2t[m][c] = pk->a2t[m][c][l];
    float acc_w2[kMlpHid];                          // dW2[o2 = l][o1]
    float acc_w1[kMlpIn];                           // dW1[o1 = l][f]
#pragma unroll
    for (int k = 0; k < kMlpHid; ++k) acc_w2[k] = 0.f;
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) acc_w1[k] = 0.f;
    float acc_b1 = 0.f, acc_b2 = 0.f, acc_w3 = 0.f, acc_b3 = 0.f, acc_loss = 0.f;
    const int ntiles = (rows + 15) / 16;
    for (int base = blockIdx.x * kWaves; base < ntiles; base += gridDim.x * kWaves) {
        const int row = (base + wv) * 16 + r;
        const bool valid = row < rows;
        const int src = valid ? (idx ? idx[row] : row) : 0;
        bf16x8 x = {};
        float tv = 0.f;
        if (valid) {
            if (g == 0) x = unpack_features(feat[src]);
            tv = targ[src];
        }
        MlpLane L;
        mlp_forward(x, R.a1, R.a2, R.b1, R.b2, R.w3, R.b3, L);
        const float err = valid ? L.y - tv : 0.f;
        const float dy = dscale * err;
        if (g == 0) acc_loss += err * err;
        float d2[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) d2[k] = L.h2[k] > 0.f ? dy * R.w3[k] : 0.f;
        bf16x8 bq[2];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int s = 0; s < 8; ++s) bq[c][s] = (__bf16)d2[8 * c + s];
        const f32x4 z = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            f32x4 d = mfma16(a2t[m][0], bq[0], z);
            d = mfma16(a2t[m][1], bq[1], d);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k = 4 * m + e, u = 16 * m + 4 * g + e;
                s_d1[wv][r][u] = L.h1[k] > 0.f ? d[e] : 0.f;
                s_h1[wv][r][u] = (float)(__bf16)L.h1[k];
                s_h2[wv][r][u] = L.h2[k];
                s_d2[wv][r][u] = d2[k];
            }
        }
        if (g == 0) {
#pragma unroll
            for (int f = 0; f < kMlpIn; ++f) s_x[wv][r][f] = (float)x[f];
            s_dy[wv][r] = dy;
        }
        __syncthreads();
        for (int q = 0; q < 16; ++q) {
            const float d2l = s_d2[wv][q][l], d1l = s_d1[wv][q][l], dyq = s_dy[wv][q];
#pragma unroll
            for (int k = 0; k < kMlpHid; ++k) acc_w2[k] += d2l * s_h1[wv][q][k];
#pragma unroll
            for (int f = 0; f < kMlpIn; ++f) acc_w1[f] += d1l * s_x[wv][q][f];
            acc_b2 += d2l;
            acc_b1 += d1l;
            acc_w3 += dyq * s_h2[wv][q][l];
            acc_b3 += dyq;
        }
        __syncthreads();
    }
    // squared errors: lanes 0-15 hold one row each
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) acc_loss += __shfl_xor(acc_loss, o, 64);
    float* P = partial + (size_t)(blockIdx.x * kWaves + wv) * kPartial;
#pragma unroll
    for (int k = 0; k < kMlpHid; ++k) P[kOffW2 + l * kMlpHid + k] = acc_w2[k];
#pragma unroll
    for (int f = 0; f < kMlpIn; ++f) P[kOffW1 + l * kMlpIn + f] = acc_w1[f];
    P[kOffB1 + l] = acc_b1;
    P[kOffB2 + l] = acc_b2;
    P[kOffW3 + l] = acc_w3;
    if (l == 0) {
        P[kOffB3] = acc_b3;
        P[kMlpParams] = acc_loss;
    }
}

// Fixed-order sum over the wavefront partials; Adam (Kingma & Ba 2015) on
// the FP32 parameters when `update`; the gradient to `grad` when non-null;
// the squared-error sum to *sq.
__global__ void k_mlp_adam(const float* __restrict__ partial, int nparts, float* __restrict__ prm,
                           float* __restrict__ m1, float* __restrict__ m2, float lr, float beta1,
                           float beta2, float eps, float c1, float c2, int update,
                           float* __restrict__ grad, float* __restrict__ sq) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > kMlpParams) return;
    float s = 0.f;
    for (int w = 0; w < nparts; ++w) s += partial[(size_t)w * kPartial + p];
    if (p == kMlpParams) {
        *sq = s;
        return;
    }
    if (grad) grad[p] = s;
    if (!update) return;
    const float a = beta1 * m1[p] + (1.f - beta1) * s;
    const float b = beta2 * m2[p] + (1.f - beta2) * s * s;
    m1[p] = a;
    m2[p] = b;
    prm[p] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

}  // namespace

int mlp_train_blocks(int rows) {
    const int tiles = (rows + 15) / 16;
    int b = (tiles + kWaves - 1) / kWaves;
    if (b > 256) b = 256;
    return b < 1 ? 1 : b;
}

size_t mlp_partial_floats(int blocks) { return (size_t)blocks * kWaves * kPartial; }

hipError_t mlp_pack(const float* prm, MlpPacked* pk, hipStream_t st) {
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, st, prm, pk);
    return hipGetLastError();
}

hipError_t mlp_rows(const uint4* feat, int rows, const MlpPacked* pk, float* out, hipStream_t st) {
    if (rows <= 0) return hipSuccess;
    const int tiles = (rows + 15) / 16;
    int blocks = (tiles + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_mlp_rows, dim3(blocks), dim3(256), 0, st, feat, rows, pk, out);
    return hipGetLastError();
}

hipError_t mlp_train_step(const uint4* feat, const float* targ, const int* idx, int rows,
                          float dscale, const MlpPacked* pk, float* partial, float* prm,
                          float* m1, float* m2, const AdamStep& a, float* grad, float* sq,
                          hipStream_t st) {
    const int blocks = mlp_train_blocks(rows);
    hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(256), 0, st, feat, targ, idx, rows, dscale,
                       pk, partial);
    hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 256) / 256), dim3(256), 0, st, partial,
                       blocks * kWaves, prm, m1, m2, a.lr, a.beta1, a.beta2, a.eps, a.c1, a.c2,
                       a.update, grad, sq);
    return hipGetLastError();
}

}  // namespace xdem
