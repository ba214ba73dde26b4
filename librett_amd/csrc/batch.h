// Batched engine: many independent transposes in one persistent launch.
//
// Every problem gets a plan from the batch-mode planner (256-thread
// workgroups, small LDS, a fixed set of kernel shapes) and a descriptor in
// one device arena. The work of all problems is a single index space of
// units (tiles, row tiles, or a few staged rounds), laid out problem after
// problem; workgroup g of the persistent kernel takes units g, g + G,
// g + 2G, ... so neighbouring units of one problem run at the same time on
// different CUs, and no counters or atomics are needed (the launch can be
// replayed from a HIP graph as is).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "planner.h"

namespace lrt {

struct BatchDesc {
  int variant;       // index into the persistent kernel's shape table
  int units;         // work units of this problem
  int rounds;        // items per unit (at most)
  int items;         // tiles (Tiled, TiledCopy) or outer indices (Packed) of the problem
  int groups;        // unit g of a split chunk takes items g, g + groups, g + 2 groups, ...
  int item_bytes;    // algorithmic bytes (read + write) of one item
  FastDiv split_div; // Packed: number of chunks of the split dim
  const void* in;
  void* out;
  union Args {
    TiledArgs tiled;
    CopyArgs copy;
    PackedArgs packed;
  } a;
};

// Shape-table index of a batch-mode candidate, or -1.
int batch_variant(const Candidate& c);

// Default algorithmic bytes of one work unit.
constexpr long long kBatchUnitBytes = 32 << 10;

// Fills d (except pointers) for a candidate; returns false if unsupported.
bool batch_describe(const Candidate& c, BatchDesc& d);

// Regroups d's items into units of about `unit_bytes` algorithmic bytes
// (at least one item per unit).
void batch_regroup(BatchDesc& d, long long unit_bytes);

// Unit size for a batch of `total_bytes` algorithmic bytes run by `grid`
// workgroups: about 32 units per workgroup, between 16 KiB and
// kBatchUnitBytes (environment overrides: LIBRETT_BATCH_UNIT_KB,
// LIBRETT_BATCH_UNITS_PER_WG).
long long batch_unit_bytes(long long total_bytes, int grid);

// Launches the persistent kernel over `total` units of `nprob` problems.
// `start` holds the first unit of every problem plus `total` at the end.
hipError_t launch_batch(const BatchDesc* d_descs, const int* d_start, int nprob, int total, int grid,
                        hipStream_t stream);

// Resident workgroups per CU of the persistent kernel (runtime query).
int batch_occupancy();

// Descriptors of one problem planned with candidate c: one descriptor, or
// up to two for a plain copy (Trivial). `fallback` gets the variants for
// buffers not aligned to the vector width; `offsets` the byte offset of each
// descriptor's data inside the problem's buffers.
bool describe_problem(const Problem& p, const Candidate& c, std::vector<BatchDesc>& descs,
                      std::vector<BatchDesc>& fallback, std::vector<long long>& offsets);

// Pointer alignment the descriptors of candidate c need.
int batch_alignment(const Problem& p, const Candidate& c);

// Batch-mode candidates of a problem, best modelled streaming cost first.
std::vector<Candidate> batch_candidates(const Problem& p, const DeviceInfo& dev);

// Host replay of the persistent kernel (descriptors hold host pointers).
// Returns false if a unit wrote past the kernel's LDS.
bool replay_batch(const BatchDesc* descs, const int* start, int nprob, int total, int grid);

}  // namespace lrt
