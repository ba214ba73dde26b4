// On-device index-pattern checker of transposes (see checker.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "common.h"

namespace lrt {

struct CheckArgs {
  int rank;
  int words;                  // 32-bit words per element
  uint32_t vol;
  FastDiv div[kMaxRank];      // output dims, fastest first
  uint32_t stride[kMaxRank];  // input stride of each output dim
};

struct CheckPartial {
  unsigned long long mismatches;
  uint32_t first;             // first mismatching output element of the workgroup, 0xffffffff if none
  uint32_t pad_;
};

constexpr int kCheckMaxGrid = 4096;

bool make_check_args(int rank, const int* dim, const int* perm, int elem, CheckArgs& a);
hipError_t launch_check_fill(void* data, size_t nbytes, hipStream_t stream);
int check_grid(const CheckArgs& a);
hipError_t launch_check(const CheckArgs& a, const void* out, CheckPartial* part, int grid, hipStream_t stream);
void host_check(const CheckArgs& a, const void* out, unsigned long long& mismatches, long long& first);

}  // namespace lrt
