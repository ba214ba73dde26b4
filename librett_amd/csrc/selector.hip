// Learned plan selector: HIP inference and training kernels of the MLP
// (selector_ops.h) and their host replays.
#include <hip/hip_runtime.h>

#include <vector>

#include "selector.h"
#include "selector_ops.h"

namespace lrt {
namespace mlp {

// Inference: one thread per sample, parameters staged in LDS.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ params, const float* __restrict__ xs,
                                                     int n, float* __restrict__ out) {
  __shared__ float p[P];
  // blockIdx.y selects the network of an ensemble.
  params += (long long)blockIdx.y * P;
  out += (long long)blockIdx.y * n;
  for (int e = threadIdx.x; e < P; e += blockDim.x) p[e] = params[e];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[F];
#pragma unroll
  for (int i = 0; i < F; i++) x[i] = xs[(long long)s * F + i];
  out[s] = forward(p, x);
}

// Training: workgroup b computes the gradient of samples
// idx[base + b*S .. base + b*S + count) into partial[b].
__global__ __launch_bounds__(T) void k_mlp_grad(const float* __restrict__ params, const float* __restrict__ xs,
                                                const float* __restrict__ ys, const float* __restrict__ ws,
                                                const int* __restrict__ idx, int base, int end, float inv_m,
                                                float* __restrict__ partial, float* __restrict__ loss) {
  extern __shared__ float lds[];
  const int first = base + blockIdx.x * S;
  const int count = max(0, min(S, end - first));
  const int* id = idx + first;
  float* grad = partial + (long long)blockIdx.x * P;
  const int t = threadIdx.x;
  phase_load(t, lds, params, xs, id, count);
  __syncthreads();
  phase_forward(t, lds, ys, ws, id, count, inv_m);
  __syncthreads();
  phase_w3(t, lds, grad, loss + blockIdx.x);
  __syncthreads();
  phase_dz2(t, lds);
  __syncthreads();
  phase_w2(t, lds, grad);
  __syncthreads();
  phase_dz1(t, lds);
  __syncthreads();
  phase_w1(t, lds, grad);
}

__global__ __launch_bounds__(256) void k_mlp_adam(const float* __restrict__ partial, int nparts,
                                                  float* __restrict__ params, float* __restrict__ m,
                                                  float* __restrict__ v, float lr, float beta1, float beta2, float c1,
                                                  float c2, float eps) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < P) adam_update(e, partial, nparts, params, m, v, lr, beta1, beta2, c1, c2, eps);
}

hipError_t launch_forward(const float* params, const float* xs, int n, float* out, hipStream_t stream,
                          int models) {
  if (n <= 0) return hipSuccess;
  if (models < 1) return hipErrorInvalidValue;
  void* args[4] = {&params, &xs, &n, &out};
  return hipLaunchKernel(reinterpret_cast<const void*>(&k_mlp_forward), dim3((n + 255) / 256, models), dim3(256),
                         args, 0, stream);
}

hipError_t prepare_grad() {
  return hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mlp_grad), hipFuncAttributeMaxDynamicSharedMemorySize,
                             kLdsFloats * (int)sizeof(float));
}

hipError_t launch_grad(const float* params, const float* xs, const float* ys, const float* ws, const int* idx,
                       int base, int end, float inv_m, int nblocks, float* partial, float* loss, hipStream_t stream) {
  void* args[10] = {&params, &xs, &ys, &ws, &idx, &base, &end, &inv_m, &partial, &loss};
  return hipLaunchKernel(reinterpret_cast<const void*>(&k_mlp_grad), dim3(nblocks), dim3(T), args,
                         kLdsFloats * sizeof(float), stream);
}

hipError_t launch_adam(const float* partial, int nparts, float* params, float* m, float* v, float lr, float beta1,
                       float beta2, float c1, float c2, float eps, hipStream_t stream) {
  void* args[11] = {&partial, &nparts, &params, &m, &v, &lr, &beta1, &beta2, &c1, &c2, &eps};
  return hipLaunchKernel(reinterpret_cast<const void*>(&k_mlp_adam), dim3((P + 255) / 256), dim3(256), args, 0,
                         stream);
}

// ---------------------------------------------------------------------------
// Host replays (same phase functions, every thread of a phase before the
// next phase).
// ---------------------------------------------------------------------------

void host_forward(const float* params, const float* xs, int n, float* out) {
  for (int s = 0; s < n; s++) out[s] = forward(params, xs + (long long)s * F);
}

void host_grad(const float* params, const float* xs, const float* ys, const float* ws, const int* idx, int count,
               float inv_m, float* grad, float* loss) {
  std::vector<float> lds(kLdsFloats);
  float* L = lds.data();
  for (int t = 0; t < T; t++) phase_load(t, L, params, xs, idx, count);
  for (int t = 0; t < T; t++) phase_forward(t, L, ys, ws, idx, count, inv_m);
  for (int t = 0; t < T; t++) phase_w3(t, L, grad, loss);
  for (int t = 0; t < T; t++) phase_dz2(t, L);
  for (int t = 0; t < T; t++) phase_w2(t, L, grad);
  for (int t = 0; t < T; t++) phase_dz1(t, L);
  for (int t = 0; t < T; t++) phase_w1(t, L, grad);
}

void host_adam(const float* partial, int nparts, float* params, float* m, float* v, float lr, float beta1,
               float beta2, float c1, float c2, float eps) {
  for (int e = 0; e < P; e++) adam_update(e, partial, nparts, params, m, v, lr, beta1, beta2, c1, c2, eps);
}

}  // namespace mlp
}  // namespace lrt
