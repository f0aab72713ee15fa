// nrc.h — neural radiance cache state and launch interface.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx_internal.h"

namespace rtx {

struct NrcState {
    float* params = nullptr;    // FP32 master parameters (RTX_NRC_PARAM_COUNT)
    float* m = nullptr;         // Adam first moment
    float* v = nullptr;         // Adam second moment
    void* fwd = nullptr;        // BF16 forward fragments
    void* bwd = nullptr;        // BF16 backward (transposed) fragments
    float* partials = nullptr;  // per-block gradients of one step
    float* loss = nullptr;      // mean loss of the last step
    float* loss_part = nullptr; // per-block losses of one step
    uint32_t max_blocks = 0;
    uint32_t cu_count = 256;    // compute units of the device (persistent inference grid)
    uint32_t batch_size = 0;
    uint32_t steps = 0;
    uint32_t cache_vertex = 1;
    uint32_t train_suffix = 0;  // training-path vertices past the cache vertex before the self-training query
    float lr = 1e-3f, beta1 = 0.9f, beta2 = 0.99f, eps = 1e-8f;
    float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
};

uint64_t nrc_fragment_bytes();
hipError_t nrc_pack(NrcState& st, hipStream_t stream);
// Optional epilogue of an inference: instead of writing the predictions (out_rgb may then be null), add
// throughput x prediction to the radiance of each query's path (query i belongs to path pid[i]; one query per
// path).
struct InferGather {
    const uint32_t* pid = nullptr;
    const float4* thr = nullptr;
    float4* rad = nullptr;
};
// radiance predictions for n queries (n_dev, when set, overrides n with a device count bounded by n)
hipError_t nrc_infer(const NrcState& st, const rtx_nrc_query* queries, uint32_t n, float* out_rgb, hipStream_t stream,
                     const uint32_t* n_dev = nullptr, const InferGather* gather = nullptr);
// one Adam step over n samples, taking sample (i * stride + offset) for i < n; with grad_out: gradient only
hipError_t nrc_step(NrcState& st, const rtx_nrc_sample* samples, uint32_t n, float* grad_out, float* loss_out,
                    hipStream_t stream, uint32_t stride = 1, uint32_t offset = 0);

}  // namespace rtx
