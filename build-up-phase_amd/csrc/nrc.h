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
    uint32_t* steps_dev = nullptr;  // Adam steps taken (device word: the bias corrections read it)
    uint32_t max_blocks = 0;
    uint32_t cu_count = 256;    // compute units of the device (persistent inference grid)
    uint32_t batch_size = 0;
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
// one Adam step over n samples (sample i of the batch is samples[i]); with grad_out: gradient only, no update
hipError_t nrc_step(NrcState& st, const rtx_nrc_sample* samples, uint32_t n, float* grad_out, float* loss_out,
                    hipStream_t stream);
// Adam steps over a device-resident sample count: batch b of ceil(count / batch_size) interleaved batches takes
// samples b, b + nb, b + 2 nb, ...; steps b >= nb return at once.  The enqueued work does not depend on the count,
// so a frame needs no host round trip (and can be captured).  Launches max_batches steps, then advances the step
// counter by nb.
hipError_t nrc_steps_device(NrcState& st, const rtx_nrc_sample* samples, const uint32_t* count_dev,
                            uint32_t max_batches, hipStream_t stream);

}  // namespace rtx
