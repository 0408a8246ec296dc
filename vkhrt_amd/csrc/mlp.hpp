// Device-side interface of the neural shadow-visibility MLP kernels (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "internal.hpp"

namespace vk {

constexpr int kMlpParamCount = 18753;

struct MlpDev {
    float* master;    // FP32 parameters (flat layout)
    float* adam_m;    // Adam first moments
    float* adam_v;    // Adam second moments
    uint16_t* packed; // bf16 MFMA fragments of W1..W4 and the transposed W2..W4
    float center[3];  // scene box centre and 1 / half extent (position normalization)
    float inv_half[3];
};

struct MlpTrainParams {
    const float4* samples;  // total x (p, 0, d, 0)
    const float* targets;   // total x visibility in {0, 1}
    uint32_t n;             // samples in this step
    uint32_t total;         // samples in the data set
    uint32_t first;         // index of the step's first sample in the epoch order
    uint64_t perm_mul, perm_add;  // epoch order i -> (i mul + add) mod total
    float* partial;         // [blocks][kMlpParamCount]
    float* loss;            // summed binary cross-entropy (device)
};

struct AdamParams {
    float lr, beta1, beta2, eps, bias1, bias2;  // bias1/2 = 1 - beta^step
};

struct MlpShadowParams {
    const uint32_t* queue;    // shadow query ids (pixel * light samples + sample); vis is indexed by id
    const uint32_t* count;    // their number (device)
    const float4* origins;    // shadow-ray origins per queue entry
    float dir[3];             // light direction
    float threshold;          // exact fallback when p (1 - p) > threshold
    float* vis;               // per-pixel visibility
    uint32_t* fallback_queue; // ids left to the exact shadow ray
    uint32_t* fallback_count;
    const float4* dirs;       // per queue entry directions (area light), null = `dir` for all
    uint32_t paired;          // 1: entries 2k, 2k + 1 are the two light samples of one hit (one origin)
};

size_t mlp_train_lds_bytes();
hipError_t mlp_enable_train_lds();
hipError_t launch_mlp_pack(const MlpDev& m, hipStream_t st);
hipError_t launch_mlp_infer(const MlpDev& m, const float* samples, uint32_t n, float* out, hipStream_t st);
hipError_t launch_mlp_shadow_gate(const MlpDev& m, const MlpShadowParams& sp, uint32_t max_count, hipStream_t st);
hipError_t launch_mlp_train_step(const MlpDev& m, const MlpTrainParams& tp, uint32_t nblocks, float* grad_out,
                                 const AdamParams& ap, hipStream_t st);
hipError_t launch_mlp_collect(const float* rays, const HitRec* hits, uint32_t n, float* samples, float* targets,
                              hipStream_t st);

}  // namespace vk
