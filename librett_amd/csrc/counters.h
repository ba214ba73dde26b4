// Exact global-memory access counts of a candidate plan, computed on the
// device (or by the same code on the host) over every work item of the
// launch. Reference counterpart: librettGpuModelKernel and its MemStat
// counters (src/GpuModelKernel.cpp:43-64, 343-946, 1013-1224).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "planner.h"

namespace lrt {

enum CountIndex {
  kLdReq = 0,     // wave load instructions with at least one active lane
  kStReq,         // wave store instructions with at least one active lane
  kLdTran,        // distinct 128-byte segments of each load instruction, summed
  kStTran,        // distinct 128-byte segments of each store instruction, summed
  kLdBytes,       // bytes loaded by active lanes
  kStBytes,       // bytes stored by active lanes
  kReadLines,     // distinct 128-byte lines each work item reads, summed over items
  kWriteLines,    // distinct 128-byte lines each work item writes, summed
  kPartialLines,  // written lines a work item covers only in part, summed
  kItems,         // work items: tiles (Tiled, TiledCopy) or staged rounds (Packed)
  kNumCounts
};

struct AccessCounts {
  unsigned long long v[kNumCounts];
};

// One global access: work item, 0 load / 1 store, instruction (pass index
// for the enumeration, call order within the thread for the kernel trace),
// thread, byte offset from the tensor start, bytes.
struct AccessRecord {
  long long item;
  int store;
  int inst;
  int thread;
  long long offset;
  int bytes;
};

// Counts on the current device; synchronises `stream`.
hipError_t device_counts(const Candidate& c, hipStream_t stream, AccessCounts& out);

// The same counts from the same per-item code run on the host.
bool host_counts(const Candidate& c, AccessCounts& out);

// Every access of the counters' enumeration (host), in item, load/store,
// instruction, thread order.
bool trace_enumeration(const Candidate& c, std::vector<AccessRecord>& recs);

// Every access the transpose kernel's host replay makes (the kernel code
// itself, run on the host), in the order it makes them.
bool trace_kernel(const Candidate& c, std::vector<AccessRecord>& recs);

}  // namespace lrt
