// Learned plan selector: features of (problem, candidate) pairs, the MLP's
// HIP kernels and host replays, and the selector state librettPlan consults.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <vector>

#include "planner.h"

namespace lrt {

namespace mlp {

// out[k * n + i] = network k (parameters params + k * P) on sample i, for
// k < models.
hipError_t launch_forward(const float* params, const float* xs, int n, float* out, hipStream_t stream,
                          int models = 1);
hipError_t prepare_grad();
hipError_t launch_grad(const float* params, const float* xs, const float* ys, const float* ws, const int* idx,
                       int base, int end, float inv_m, int nblocks, float* partial, float* loss, hipStream_t stream);
hipError_t launch_adam(const float* partial, int nparts, float* params, float* m, float* v, float lr, float beta1,
                       float beta2, float c1, float c2, float eps, hipStream_t stream);

void host_forward(const float* params, const float* xs, int n, float* out);
void host_grad(const float* params, const float* xs, const float* ys, const float* ws, const int* idx, int count,
               float inv_m, float* grad, float* loss);
void host_adam(const float* partial, int nparts, float* params, float* m, float* v, float lr, float beta1,
               float beta2, float c1, float c2, float eps);

}  // namespace mlp

// Largest reduced rank the features describe; larger problems use the
// analytic model.
constexpr int kSelMaxRank = 7;

// Features per (problem, candidate) pair (the MLP's input width).
constexpr int kSelFeatures = 50;
// Largest ensemble the selector averages.
constexpr int kSelMaxModels = 16;

// Raw (unnormalised) features of candidate c of problem `full`; `reduced`
// is reduce_ranks(full). Writes mlp::F floats.
void selector_features(const Problem& full, const Problem& reduced, const Candidate& c, float* f);

// Index of the candidate the learned selector picks, or -1 when no weights
// are loaded, the problem is outside the features' domain or inference
// fails (the caller then uses the analytic model). Runs the HIP inference
// kernel on the current device.
int selector_choose(const std::vector<Candidate>& cands, const Problem& full, int device);

// Loads the weights (if any) and makes the device's inference buffers now
// rather than in the first plan.
void selector_prepare(int device);

// Returns the per-device inference buffers (taken from the device allocator,
// kept between plans) to the allocator that made them, then runs `then`
// with the selector locked, so no plan can take a buffer in between. The
// allocator hook calls it when the allocator changes; the buffers are made
// again from the new allocator on the next plan.
void selector_release_buffers(const std::function<void()>& then);

// Generation of the loaded weights (0: none); part of the plan-cache key.
uint64_t selector_generation();

// The selector overrides the analytic model's choice only when its scores
// (predicted log2 time over the roofline time) put its own pick below the
// analytic choice by more than this margin. 0.03 (about 2%) by default;
// LIBRETT_SELECTOR_MARGIN sets another value (0: the selector's pick
// always).
constexpr double kSelMargin = 0.03;
double selector_margin();

}  // namespace lrt
