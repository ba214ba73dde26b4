// Launch, replay and occupancy entry points of the kernel table (kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "planner.h"

namespace lrt {

// Enqueues the candidate's kernel (or the D2D copy for Trivial) on `stream`.
hipError_t launch_candidate(const Candidate& c, const void* in, void* out, hipStream_t stream);

// Runs the candidate's per-thread code on the host (CPU tests). Returns
// false if the candidate names a kernel shape that is not instantiated.
bool replay_candidate(const Candidate& c, const void* in, void* out);

// Runtime occupancy (active blocks per CU) of the candidate's kernel.
int query_occupancy(const Candidate& c);

// Sets per-kernel attributes a launch needs (dynamic LDS size).
int prepare_kernel(const Candidate& c);

// Device-to-device copy kernel (bandwidth ceiling): k 16-byte words per
// thread per round, nt != 0 for non-temporal stores, grid <= 0 for one block
// per 256*k words.
hipError_t launch_device_copy(void* dst, const void* src, size_t bytes, hipStream_t stream, int k, int nt, int grid);

// Loads the transpose kernels' code object on the current device (the
// runtime does it on first use, which would otherwise land in the first
// plan's occupancy queries).
void kernels_warm_up();

// general: a kernel for planes whose x or y is a group of several dims.
bool tiled_shape_available(int elem, int tx, int ty, int vx, int vy, bool general = false);
// Shapes the batched engine's persistent kernel is instantiated for (batch.hip).
bool batch_tiled_available(int elem, int tx, int ty, int vx, int vy, bool general = false);
bool batch_copy_available(int word, int tx);
bool batch_packed_available(int elem, int nreg);
bool copy_shape_available(int word, int tx);
int packed_max_nreg(int elem);
bool packed_nreg_available(int elem, int nreg);

}  // namespace lrt
