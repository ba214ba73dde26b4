// Learned plan selector: the 2 x 64 FP32 MLP that scores candidate plans,
// written as per-thread phase functions shared by the HIP kernels
// (selector.hip) and their host replay. A workgroup runs the phases in
// order with a barrier between them; the host replay runs each phase for
// every thread before starting the next one.
//
// The reference has no learned selector; it ranks candidates with the
// analytic GPU model only (src/plan.cpp:721-789, src/GpuModel.cpp:798-911).
#pragma once

#include <hip/hip_runtime.h>

namespace lrt {
namespace mlp {

constexpr int F = 50;  // features per (problem, candidate)
constexpr int H = 64;  // width of both hidden layers

// Parameter layout (floats): W1[H][F], b1[H], W2[H][H], b2[H], w3[H], b3.
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + H * F;
constexpr int kW2 = kB1 + H;
constexpr int kB2 = kW2 + H * H;
constexpr int kW3 = kB2 + H;
constexpr int kB3 = kW3 + H;
constexpr int P = kB3 + 1;

// Training: samples per workgroup and threads per workgroup.
constexpr int S = 128;
constexpr int T = 256;

// LDS of the training workgroup (floats): parameters, inputs, two layers of
// pre-activations (overwritten by their gradients), per-sample loss
// gradients and squared errors.
constexpr int kLdsP = 0;
constexpr int kLdsX = kLdsP + P;
constexpr int kLdsZ1 = kLdsX + S * F;
constexpr int kLdsZ2 = kLdsZ1 + S * H;
constexpr int kLdsD = kLdsZ2 + S * H;
constexpr int kLdsE = kLdsD + S;
constexpr int kLdsFloats = kLdsE + S;

__host__ __device__ inline float relu(float v) { return v > 0.0f ? v : 0.0f; }

// Forward pass of one sample; x has F normalised features.
__host__ __device__ inline float forward(const float* __restrict__ p, const float* __restrict__ x) {
  float a1[H];
#pragma unroll
  for (int j = 0; j < H; j++) {
    float z = p[kB1 + j];
#pragma unroll
    for (int i = 0; i < F; i++) z += p[kW1 + j * F + i] * x[i];
    a1[j] = relu(z);
  }
  float out = p[kB3];
  for (int j = 0; j < H; j++) {
    float z = p[kB2 + j];
#pragma unroll
    for (int k = 0; k < H; k++) z += p[kW2 + j * H + k] * a1[k];
    out += p[kW3 + j] * relu(z);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Training phases. One workgroup handles `count` (<= S) samples whose rows
// are idx[0..count); the loss is sum w (out - y)^2 * inv_m over the
// minibatch (w = 1 when no weights are given; inv_m = 1 / sum w), so the
// workgroup's gradient is its share of the minibatch gradient.
// ---------------------------------------------------------------------------

// Phase 0: parameters and the workgroup's inputs into LDS.
__host__ __device__ inline void phase_load(int tid, float* lds, const float* __restrict__ params,
                                           const float* __restrict__ xs, const int* __restrict__ idx, int count) {
  for (int e = tid; e < P; e += T) lds[kLdsP + e] = params[e];
  for (int e = tid; e < S * F; e += T) {
    const int s = e / F, i = e - (e / F) * F;
    lds[kLdsX + e] = s < count ? xs[(long long)idx[s] * F + i] : 0.0f;
  }
}

// Phase 1: forward pass per sample; keeps pre-activations, d loss / d out
// and the squared error.
__host__ __device__ inline void phase_forward(int tid, float* lds, const float* __restrict__ ys,
                                              const float* __restrict__ ws, const int* __restrict__ idx, int count,
                                              float inv_m) {
  const float* p = lds + kLdsP;
  for (int s = tid; s < S; s += T) {
    float* z1 = lds + kLdsZ1 + s * H;
    float* z2 = lds + kLdsZ2 + s * H;
    if (s >= count) {
      for (int j = 0; j < H; j++) z1[j] = z2[j] = 0.0f;
      lds[kLdsD + s] = 0.0f;
      lds[kLdsE + s] = 0.0f;
      continue;
    }
    const float* x = lds + kLdsX + s * F;
    for (int j = 0; j < H; j++) {
      float z = p[kB1 + j];
      for (int i = 0; i < F; i++) z += p[kW1 + j * F + i] * x[i];
      z1[j] = z;
    }
    float out = p[kB3];
    for (int j = 0; j < H; j++) {
      float z = p[kB2 + j];
      for (int k = 0; k < H; k++) z += p[kW2 + j * H + k] * relu(z1[k]);
      z2[j] = z;
      out += p[kW3 + j] * relu(z);
    }
    const float err = out - ys[idx[s]];
    const float w = ws ? ws[idx[s]] : 1.0f;
    lds[kLdsD + s] = 2.0f * w * err * inv_m;
    lds[kLdsE + s] = w * err * err;
  }
}

// Phase 2: gradients of w3 and b3, and the workgroup's weighted squared-error sum.
__host__ __device__ inline void phase_w3(int tid, const float* lds, float* __restrict__ grad,
                                         float* __restrict__ loss) {
  for (int j = tid; j < H + 2; j += T) {
    float g = 0.0f;
    if (j < H) {
      for (int s = 0; s < S; s++) g += lds[kLdsD + s] * relu(lds[kLdsZ2 + s * H + j]);
      grad[kW3 + j] = g;
    } else if (j == H) {
      for (int s = 0; s < S; s++) g += lds[kLdsD + s];
      grad[kB3] = g;
    } else {
      for (int s = 0; s < S; s++) g += lds[kLdsE + s];
      *loss = g;
    }
  }
}

// Phase 3: Z2 <- d loss / d z2.
__host__ __device__ inline void phase_dz2(int tid, float* lds) {
  for (int e = tid; e < S * H; e += T) {
    const int s = e / H, j = e - (e / H) * H;
    const float z = lds[kLdsZ2 + e];
    lds[kLdsZ2 + e] = z > 0.0f ? lds[kLdsD + s] * lds[kLdsP + kW3 + j] : 0.0f;
  }
}

// Phase 4: gradients of W2 and b2.
__host__ __device__ inline void phase_w2(int tid, const float* lds, float* __restrict__ grad) {
  for (int e = tid; e < H * H + H; e += T) {
    float g = 0.0f;
    if (e < H * H) {
      const int j = e / H, k = e - (e / H) * H;
      for (int s = 0; s < S; s++) g += lds[kLdsZ2 + s * H + j] * relu(lds[kLdsZ1 + s * H + k]);
      grad[kW2 + e] = g;
    } else {
      const int j = e - H * H;
      for (int s = 0; s < S; s++) g += lds[kLdsZ2 + s * H + j];
      grad[kB2 + j] = g;
    }
  }
}

// Phase 5: Z1 <- d loss / d z1 (each element read and written by one thread).
__host__ __device__ inline void phase_dz1(int tid, float* lds) {
  for (int e = tid; e < S * H; e += T) {
    const int s = e / H, k = e - (e / H) * H;
    float g = 0.0f;
    if (lds[kLdsZ1 + e] > 0.0f)
      for (int j = 0; j < H; j++) g += lds[kLdsP + kW2 + j * H + k] * lds[kLdsZ2 + s * H + j];
    lds[kLdsZ1 + e] = g;
  }
}

// Phase 6: gradients of W1 and b1.
__host__ __device__ inline void phase_w1(int tid, const float* lds, float* __restrict__ grad) {
  for (int e = tid; e < H * F + H; e += T) {
    float g = 0.0f;
    if (e < H * F) {
      const int j = e / F, i = e - (e / F) * F;
      for (int s = 0; s < S; s++) g += lds[kLdsZ1 + s * H + j] * lds[kLdsX + s * F + i];
      grad[kW1 + e] = g;
    } else {
      const int j = e - H * F;
      for (int s = 0; s < S; s++) g += lds[kLdsZ1 + s * H + j];
      grad[kB1 + j] = g;
    }
  }
}

// Adam step for parameter e: the gradient is the sum of the workgroups'
// partial gradients in workgroup order. c1 = 1 / (1 - beta1^t) and
// c2 = 1 / (1 - beta2^t) are the bias corrections of step t.
__host__ __device__ inline void adam_update(int e, const float* __restrict__ partial, int nparts,
                                            float* __restrict__ params, float* __restrict__ m,
                                            float* __restrict__ v, float lr, float beta1, float beta2, float c1,
                                            float c2, float eps) {
  float g = 0.0f;
  for (int b = 0; b < nparts; b++) g += partial[(long long)b * P + e];
  const float mm = beta1 * m[e] + (1.0f - beta1) * g;
  const float vv = beta2 * v[e] + (1.0f - beta2) * g * g;
  m[e] = mm;
  v[e] = vv;
  params[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

}  // namespace mlp
}  // namespace lrt
