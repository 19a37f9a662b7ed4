// mlp_launch.h -- host launchers of the root-guess network kernels
// (mlp_kernels.hip), used by the engine's C-ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace xdem {

struct MlpPacked;

struct AdamStep {
    float lr, beta1, beta2, eps;
    float c1, c2;            // 1 / (1 - beta1^t), 1 / (1 - beta2^t)
    int update;              // 0: gradient and loss only
};

// Training blocks for a batch of `rows` rows and the size of the partial
// buffer they fill (floats).
int mlp_train_blocks(int rows);
size_t mlp_partial_floats(int blocks);

hipError_t mlp_pack(const float* prm, MlpPacked* pk, hipStream_t st);
hipError_t mlp_rows(const uint4* feat, int rows, const MlpPacked* pk, float* out, hipStream_t st);
// Forward/backward over rows idx[0..rows) (idx == nullptr: 0..rows), then the
// reduction and (a.update) the Adam step; grad (optional) receives
// d loss / d params, *sq the sum of squared errors.
hipError_t mlp_train_step(const uint4* feat, const float* targ, const int* idx, int rows,
                          float dscale, const MlpPacked* pk, float* partial, float* prm,
                          float* m1, float* m2, const AdamStep& a, float* grad, float* sq,
                          hipStream_t st);

}  // namespace xdem
