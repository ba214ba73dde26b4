// On-device check of a transpose against an index pattern, for outputs too
// large to copy back and compare on the host. The reference's counterpart is
// TensorTester (src/TensorTester.cpp:34-126 fills 32-bit words with their
// index and checks each output element against its source index;
// :201-220 builds the index map).
//
// Pattern: 32-bit word k of the input holds k (mod 2^32), so input element e
// of B bytes holds the words e*B/4 + w, w < B/4. The output element at
// column-major position j of dims (dim[perm[0]], dim[perm[1]], ...) must hold
// the words of input element e(j) = sum_k digit_k(j) * instride[perm[k]].
// Each workgroup writes its mismatch count and first mismatching position to
// its own slot (no atomics); the host reduces the slots.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "devmem.h"
#include "../../include/librett_c.h"
#include "checker.h"

namespace lrt {

__host__ __device__ inline bool check_element(const CheckArgs& a, uint32_t j, const uint32_t* out) {
  uint32_t k = j, e = 0;
  for (int r = 0; r < a.rank; r++) {
    const uint32_t q = fdiv(k, a.div[r]);
    e += (k - q * a.div[r].d) * a.stride[r];
    k = q;
  }
  const uint32_t base = e * (uint32_t)a.words;
  bool ok = true;
  for (int w = 0; w < a.words; w++) ok = ok && out[(size_t)j * a.words + w] == base + (uint32_t)w;
  return ok;
}

namespace {

__global__ __launch_bounds__(256) void k_check_fill(uint32_t* data, unsigned long long nwords) {
  for (unsigned long long i = blockIdx.x * 256ull + threadIdx.x; i < nwords; i += (unsigned long long)gridDim.x * 256)
    data[i] = (uint32_t)i;
}

__global__ __launch_bounds__(256) void k_check(CheckArgs a, const uint32_t* __restrict__ out,
                                               CheckPartial* __restrict__ part) {
  __shared__ unsigned long long cnt[256];
  __shared__ uint32_t first[256];
  unsigned long long c = 0;
  uint32_t f = 0xffffffffu;
  for (uint32_t j = blockIdx.x * 256u + threadIdx.x; j < a.vol; j += gridDim.x * 256u) {
    if (!check_element(a, j, out)) {
      c++;
      f = min(f, j);
    }
  }
  cnt[threadIdx.x] = c;
  first[threadIdx.x] = f;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      cnt[threadIdx.x] += cnt[threadIdx.x + s];
      first[threadIdx.x] = min(first[threadIdx.x], first[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x].mismatches = cnt[0];
    part[blockIdx.x].first = first[0];
  }
}

// Sums the per-workgroup partials of k_check into result[0] (wrong
// elements) and result[1] (first wrong output position, ~0 if none).
__global__ __launch_bounds__(256) void k_check_reduce(const CheckPartial* __restrict__ part, int n,
                                                      unsigned long long* __restrict__ result) {
  __shared__ unsigned long long cnt[256];
  __shared__ uint32_t first[256];
  unsigned long long c = 0;
  uint32_t f = 0xffffffffu;
  for (int i = threadIdx.x; i < n; i += 256) {
    c += part[i].mismatches;
    f = min(f, part[i].first);
  }
  cnt[threadIdx.x] = c;
  first[threadIdx.x] = f;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      cnt[threadIdx.x] += cnt[threadIdx.x + s];
      first[threadIdx.x] = min(first[threadIdx.x], first[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    result[0] = cnt[0];
    result[1] = cnt[0] ? (unsigned long long)first[0] : ~0ull;
  }
}

}  // namespace

// Scratch of the allocation-free checker: one partial per workgroup of
// k_check (whose grid never exceeds kCheckMaxGrid), 64 KiB. Every GPU the
// library runs on holds its own copy until the process ends. Host code
// never passes this name to a kernel or a copy: it uses the device address
// from hipGetSymbolAddress (scratch_address below).
static __device__ __align__(256) CheckPartial scratch[kCheckMaxGrid];

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept; nullptr if the lookup fails.
static CheckPartial* scratch_address() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static CheckPartial* addr[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    (void)hipGetLastError();
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    addr[dev] = static_cast<CheckPartial*>(p);
  }
  return addr[dev];
}

bool make_check_args(int rank, const int* dim, const int* perm, int elem, CheckArgs& a) {
  if (rank < 1 || rank > kMaxRank || (elem != 4 && elem != 8 && elem != 16)) return false;
  bool seen[kMaxRank] = {};
  for (int i = 0; i < rank; i++) {
    if (perm[i] < 0 || perm[i] >= rank || seen[perm[i]]) return false;
    seen[perm[i]] = true;
  }
  long long instride[kMaxRank];
  long long vol = 1;
  for (int i = 0; i < rank; i++) {
    if (dim[i] < 1) return false;
    instride[i] = vol;
    vol *= dim[i];
    if (vol >= (1LL << 31)) return false;
  }
  a.rank = rank;
  a.vol = (uint32_t)vol;
  a.words = elem / 4;
  for (int k = 0; k < rank; k++) {
    a.div[k] = make_fastdiv((uint32_t)dim[perm[k]]);
    a.stride[k] = (uint32_t)instride[perm[k]];
  }
  return true;
}

hipError_t launch_check_fill(void* data, size_t nbytes, hipStream_t stream) {
  const unsigned long long nwords = nbytes / 4;
  if (nwords == 0) return hipSuccess;
  const int grid = (int)std::min<unsigned long long>(65536, (nwords + 255) / 256);
  uint32_t* p = static_cast<uint32_t*>(data);
  void* args[2] = {&p, (void*)&nwords};
  return hipLaunchKernel(reinterpret_cast<const void*>(&k_check_fill), dim3(grid), dim3(256), args, 0, stream);
}

hipError_t launch_check(const CheckArgs& a, const void* out, CheckPartial* part, int grid, hipStream_t stream) {
  const uint32_t* p = static_cast<const uint32_t*>(out);
  CheckArgs aa = a;
  void* args[3] = {&aa, &p, &part};
  return hipLaunchKernel(reinterpret_cast<const void*>(&k_check), dim3(grid), dim3(256), args, 0, stream);
}

int check_grid(const CheckArgs& a) {
  return (int)std::max<uint32_t>(1, std::min<uint32_t>(kCheckMaxGrid, (a.vol + 255) / 256));
}

void host_check(const CheckArgs& a, const void* out, unsigned long long& mismatches, long long& first) {
  const uint32_t* p = static_cast<const uint32_t*>(out);
  mismatches = 0;
  first = -1;
  for (uint32_t j = 0; j < a.vol; j++)
    if (!check_element(a, j, p)) {
      if (first < 0) first = j;
      mismatches++;
    }
}

}  // namespace lrt

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

using namespace lrt;

extern "C" {

int librett_c_check_fill(void* data, size_t nbytes, void* stream) {
  if (!data || nbytes % 4) return LIBRETT_C_INVALID_PARAMETER;
  if (launch_check_fill(data, nbytes, (hipStream_t)stream) != hipSuccess) {
    (void)hipGetLastError();
    return LIBRETT_C_INTERNAL_ERROR;
  }
  return LIBRETT_C_SUCCESS;
}

int librett_c_check_transpose(const void* odata, int rank, const int* dim, const int* permutation, size_t sizeofType,
                              void* stream, unsigned long long* mismatches, long long* first) {
  if (!odata || !dim || !permutation || !mismatches) return LIBRETT_C_INVALID_PARAMETER;
  CheckArgs a;
  if (!make_check_args(rank, dim, permutation, (int)sizeofType, a)) return LIBRETT_C_INVALID_PARAMETER;
  const int grid = check_grid(a);
  CheckPartial* d_part = nullptr;
  std::vector<CheckPartial> h(grid);
  hipStream_t s = (hipStream_t)stream;
  bool ok = dev_malloc((void**)&d_part, grid * sizeof(CheckPartial)) == hipSuccess &&
            launch_check(a, odata, d_part, grid, s) == hipSuccess &&
            hipMemcpyAsync(h.data(), d_part, grid * sizeof(CheckPartial), hipMemcpyDeviceToHost, s) == hipSuccess &&
            hipStreamSynchronize(s) == hipSuccess;
  if (d_part) (void)dev_free(d_part);
  if (!ok) {
    (void)hipGetLastError();
    return LIBRETT_C_INTERNAL_ERROR;
  }
  unsigned long long n = 0;
  uint32_t f = 0xffffffffu;
  for (const auto& p : h) {
    n += p.mismatches;
    f = std::min(f, p.first);
  }
  *mismatches = n;
  if (first) *first = n ? (long long)f : -1;
  return LIBRETT_C_SUCCESS;
}

int librett_c_check_transpose_noalloc(const void* odata, int rank, const int* dim, const int* permutation,
                                      size_t sizeofType, void* stream, unsigned long long* result) {
  if (!odata || !dim || !permutation || !result) return LIBRETT_C_INVALID_PARAMETER;
  CheckArgs a;
  if (!make_check_args(rank, dim, permutation, (int)sizeofType, a)) return LIBRETT_C_INVALID_PARAMETER;
  const int grid = check_grid(a);
  // The scratch holds kCheckMaxGrid partials; check_grid never asks for more.
  if (grid < 1 || grid > kCheckMaxGrid) return LIBRETT_C_INVALID_PARAMETER;
  CheckPartial* part = scratch_address();
  if (!part) return LIBRETT_C_INTERNAL_ERROR;
  hipStream_t s = (hipStream_t)stream;
  // Every workgroup of k_check writes its own slot, and the reduction reads
  // only those slots, so nothing left from an earlier call is read.
  if (launch_check(a, odata, part, grid, s) != hipSuccess) {
    (void)hipGetLastError();
    return LIBRETT_C_INTERNAL_ERROR;
  }
  const CheckPartial* cpart = part;
  int n = grid;
  void* args[3] = {&cpart, &n, &result};
  if (hipLaunchKernel(reinterpret_cast<const void*>(&k_check_reduce), dim3(1), dim3(256), args, 0, s) != hipSuccess) {
    (void)hipGetLastError();
    return LIBRETT_C_INTERNAL_ERROR;
  }
  return LIBRETT_C_SUCCESS;
}

void librett_c_host_check_fill(void* data, size_t nbytes) {
  uint32_t* p = static_cast<uint32_t*>(data);
  for (size_t i = 0; i < nbytes / 4; i++) p[i] = (uint32_t)i;
}

int librett_c_host_check_transpose(const void* odata, int rank, const int* dim, const int* permutation,
                                   size_t sizeofType, unsigned long long* mismatches, long long* first) {
  if (!odata || !dim || !permutation || !mismatches) return LIBRETT_C_INVALID_PARAMETER;
  CheckArgs a;
  if (!make_check_args(rank, dim, permutation, (int)sizeofType, a)) return LIBRETT_C_INVALID_PARAMETER;
  long long f = -1;
  host_check(a, odata, *mismatches, f);
  if (first) *first = f;
  return LIBRETT_C_SUCCESS;
}

}  // extern "C"
