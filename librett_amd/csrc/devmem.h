// Device memory of the library's own buffers (batch descriptors, autotuner
// scratch, the serving arena and slots, training and checker buffers): from
// the caller's allocator when one is installed (librett_c_set_allocator),
// else hipMalloc. The reference routes its device allocations through an
// Umpire allocator chosen at librettInitialize (src/librett.cpp:387-395,
// src/GpuMem.hpp:43-83) for the same purpose: sharing a host code's pool.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace lrt {

// Allocates `bytes` of device memory; each pointer is later released by the
// allocator that made it, even if another one is installed meanwhile.
hipError_t dev_malloc(void** p, size_t bytes);
void dev_free(void* p);

}  // namespace lrt
