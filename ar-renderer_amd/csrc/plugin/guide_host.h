// Host side of learned path guiding: the model's device buffers and the launches that evaluate
// and train it (guide.hip).
#pragma once

#include <vector>

#include "device.h"
#include "guide.h"

namespace pgpu {

// The box vertex positions are normalised to before encoding.
struct GuideBox {
    float lo[3], inv[3];
};

struct GuideState {
    DevBuf params, m, v;   // FP32 master parameters and Adam moments
    DevBuf wt, bias;       // BF16 weights (Wt[out][in] per layer) and FP32 biases
    DevBuf grad, stats, part;
    DevBuf out, rec;       // per batch: raw outputs per guided vertex, training record per path
    DevBuf train, trainCount;  // compacted training records; [0] written, [1] produced
    size_t cap = 0, trainCap = 0;
    int steps = 0;
    double recordsPerPath = 0;  // training records the last training batch produced per path
    size_t ringHead = 0, ringValid = 0;  // training-record ring: next write position, records held
    size_t batchAdd = 0;                 // records the current batch may add to the ring
    size_t rotor = 0;                    // next slice of the ring to train on
    float blend = kGBlend;      // weight of the MIS-proposal term in the training loss

    void Init(hipStream_t s);
    void SetParams(const float *P, hipStream_t s);  // also resets the Adam state
    void EnsureBatch(size_t n);
    void EnsureTrain(size_t n);
    // Raw outputs for the vertices of `queue` (count on the device, at most maxCount).
    void Infer(const int *queue, const int *count, int maxCount, const float4 *h0, const float4 *h1,
               const float4 *dir, const GuideBox &box, int cus, hipStream_t s);
    // Gradient of the normalised loss over records [first, first + n) (clipped to *count when
    // count is given) into grad; stats = {summed weight, summed weighted loss}.
    void Gradient(const float *recs, const int *count, int first, int n, hipStream_t s);
    // Gradient, then one Adam step with learning rate lr.
    void Step(const float *recs, const int *count, int first, int n, float lr, hipStream_t s);
    void Eval(const float *recs9, int n, float *out41, hipStream_t s);
};

std::vector<float> GuideInitialParams();
size_t GuideTrainLdsBytes();

}  // namespace pgpu
