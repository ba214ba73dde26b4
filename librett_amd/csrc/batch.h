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

// The persistent kernel comes in one instantiation per family of shapes, so
// that each holds only the registers its own shapes need.
enum BatchFamily : int { kFamTiled = 0, kFamCopy = 1, kFamPacked = 2, kFamTiledG = 3 };
constexpr int kBatchFamilies = 4;

// Family of a shape-table index.
int batch_family(int variant);

// Descriptors grouped by family: family f owns descriptors
// [first[f], first[f] + n[f]) and start[start_off[f] ...] holds the first
// unit of each of them and then the family's unit total.
struct BatchLayout {
  int n[kBatchFamilies] = {};
  int first[kBatchFamilies] = {};
  long long total[kBatchFamilies] = {};
  int start_off[kBatchFamilies] = {};
  std::vector<int> start;  // descriptors + kBatchFamilies entries
};

// Orders `descs` by family (stable) and fills the layout; order[k] is the
// original index of the k-th ordered descriptor.
void batch_layout(std::vector<BatchDesc>& descs, BatchLayout& layout, std::vector<int>* order);

// Launches family `fam`'s persistent kernel over `total` units of `nprob`
// descriptors (`start`: first unit of each, then `total`), with at most
// `grid` workgroups.
hipError_t launch_batch(int fam, const BatchDesc* d_descs, const int* d_start, int nprob, int total, int grid,
                        hipStream_t stream);

// One launch per non-empty family, in family order on one stream; grid[f]
// caps family f's workgroups.
hipError_t launch_batch_all(const BatchDesc* d_descs, const int* d_start, const BatchLayout& layout, const int* grid,
                            hipStream_t stream);

// Resident workgroups per CU of family `fam`'s kernel (runtime query).
int batch_occupancy(int fam);

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

// Host replay of every family of a layout, `grid` workgroups each.
bool replay_batch_all(const BatchDesc* descs, const BatchLayout& layout, int grid);

}  // namespace lrt
