// Host planner: problem canonicalisation, rank reduction, candidate plans,
// kernel arguments and the gfx950 cost model.
//
// Reference counterparts: reduceRanks (src/plan.cpp:64-127), TensorSplit
// (src/plan.h:60-118, src/plan.cpp:173-481), candidate generation
// (src/plan.cpp:486-719), librettPlan_t::setup (src/plan.cpp:837-1066),
// librettKernelLaunchConfiguration (src/kernel.cpp:827-1006) and the
// countCycles model (src/plan.cpp:1073-1643, src/GpuModel.cpp:798-911).
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "common.h"

namespace lrt {

// What the planner needs to know about the device. Filled from
// hipGetDeviceProperties at run time, or by hand for host-only planning.
struct DeviceInfo {
  int device_id = 0;
  int cu_count = 256;                // MI355X
  int lds_per_block = 160 * 1024;    // gfx950: 160 KiB per workgroup
  int lds_per_cu = 160 * 1024;
  int max_threads_per_block = 1024;
  int warp_size = 64;
  int max_waves_per_cu = 32;
  double hbm_bytes_per_s = 8.0e12;   // published MI355X HBM3E peak
  double l2_bytes = 32.0 * 1024 * 1024;  // L2 of all XCDs (8 x 4 MiB on MI355X)
  int has_gpu = 0;                   // occupancy may be queried from the runtime
  // Batched engine: candidates must run in its persistent kernel (256-thread
  // workgroups, at most kBatchLds bytes of LDS, its instantiated shapes).
  int batch_mode = 0;
};

// LDS of the batched engine's persistent kernel.
constexpr int kBatchLds = 17 * 1024;

// A transpose problem: column-major input dims, output dims dim[perm[i]].
struct Problem {
  int rank = 0;
  int dim[kMaxRank];
  int perm[kMaxRank];
  int elem = 0;
  long long vol = 0;
};

// Removes size-1 dims (they change neither input nor output layout).
Problem canonicalize(int rank, const int* dim, const int* perm, int elem);

// Merges runs of dims that stay adjacent and in order in the output.
// Follows src/plan.cpp:64-127: consecutive output positions i-1, i with
// perm[i] == perm[i-1] + 1 are fused, then the surviving dims are renumbered
// and put back in input order.
Problem reduce_ranks(const Problem& p);

// Reference TensorSplit (src/plan.h:60-118): Mm = first sizeMm input dims,
// Mk = first sizeMk output dims, Mmk = Mm u Mk, Mbar = the rest.
struct Split {
  int sizeMm = 0, volMm = 1;
  int sizeMk = 0, volMk = 1;
  int sizeMmk = 0, volMmk = 1;
  int sizeMkBar = 0, volMkBar = 1;
  int sizeMbar = 0, volMbar = 1;
  int volMmkInCont = 1, volMmkOutCont = 1;
  int numSplit = 1, splitRank = -1, splitDim = 0, volMmkUnsplit = 0;
};

// Fills a Split the way TensorSplit::update does (src/plan.cpp:202-281).
void split_update(Split& s, int sizeMm, int sizeMk, const Problem& p);

// Modelled cost and the features it was computed from.
struct ModelStats {
  double read_lines = 0;      // distinct 128-byte lines read (per block window)
  double write_lines = 0;     // distinct 128-byte lines written
  double partial_lines = 0;   // written lines a block window covers only in part
  double read_tran = 0;       // 128-byte transactions of all load wave instructions
  double write_tran = 0;      // 128-byte transactions of all store wave instructions
  double items = 0;           // block rounds (tiles or staged sub-tensors)
  double concurrency = 0;     // blocks resident on the device at once
  double steps_per_block = 0; // serial load/store rounds per resident block
  double t_hbm = 0, t_latency = 0, t_launch = 0;
  double t_stream = 0;        // bandwidth-only time (lines and requests): cost inside a busy batch
  double predicted_s = 0;
};

struct Candidate {
  Method method = kUnknown;
  int elem = 0;            // element bytes
  const char* source = ""; // "reduced" or "full"
  // Tiled: tile, lane widths, and the dims in the plane's x and y groups
  int tile_x = 0, tile_y = 0, vec_x = 1, vec_y = 1;
  int xgroup = 1, ygroup = 1;
  // TiledCopy
  int word = 0, copy_tx = 0;
  // Packed / PackedSplit
  int nreg = 0;
  Split split;
  // launch
  int grid_x = 1, grid_y = 1, block = 1;
  int lds_bytes = 0;
  int active_blocks_per_cu = 0;
  long long bytes = 0;     // 2 * vol * elem
  // kernel arguments
  TiledArgs tiled;
  CopyArgs copy;
  PackedArgs packed;
  ModelStats model;
  std::string describe() const;
};

// Tile order of a Tiled or TiledCopy candidate: 1 = XCD-contiguous, 0 =
// dispatch order (0 for other methods).
inline int tile_order(const Candidate& c) {
  return (c.method == kTiled && c.tiled.xcds > 1) || (c.method == kTiledCopy && c.copy.xcds > 1) ? 1 : 0;
}

// All candidates for a problem. The list is empty only for invalid input.
// Follows the method rules of createPlans (src/plan.cpp:702-719): Trivial
// alone when the reduced rank is 1; TiledCopy when the reduced perm keeps
// dim 0 first, Tiled otherwise; Packed / PackedSplit for every (Mm, Mk)
// split of the full and the reduced problem that can be staged on chip.
std::vector<Candidate> make_candidates(const Problem& full, const DeviceInfo& dev);

// Applies the gfx950 cost model to a candidate (fills cand.model).
void model_candidate(Candidate& c, const Problem& full, const DeviceInfo& dev);

// Recomputes the modelled times from the counts already in cand.model (for
// counts replaced by exact ones).
void model_times(Candidate& c, const DeviceInfo& dev);

// Overrides the model constants (bw_frac, write_weight, partial_weight,
// inflight_half, round_latency, launch); entries <= 0 keep their value.
void set_model_constants(const double* v, int n);

// Index of the candidate with the smallest modelled time (Trivial wins, as
// in operator> of src/plan.cpp:721-768).
int choose_by_model(const std::vector<Candidate>& cands);

// Sets the Tiled tile order: groups of `group` tile rows, y fastest inside a
// group (group 1 = x fastest).
void set_tile_group(TiledArgs& a, int nty, int group);

// Variant of a Tiled candidate without vector accesses, used when the
// caller's pointers are not aligned to the vector width.
Candidate scalar_fallback(const Candidate& c);

// Pointer alignment a candidate's vector accesses need.
int needed_alignment(const Candidate& c);

// Variant of a TiledCopy candidate that moves element-sized words.
Candidate copy_fallback(const Candidate& c);

// The candidate to launch when the pointers are not aligned to
// needed_alignment(c): scalar Tiled lanes, element-sized TiledCopy words,
// or c itself (Packed and Trivial need only element alignment).
Candidate unaligned_fallback(const Candidate& c);

// Contiguous runs (start offset, length in elements) of one staged shape
// walked in LDS order (input side) or output order (output side).
void staged_runs(const StagedShape& s, bool output_side, std::vector<std::pair<long long, int>>& runs);

// Starts the planning threads now rather than in the first plan.
void planner_warm_up();

// Active blocks per CU for a candidate's kernel (runtime occupancy query if
// a GPU is present, otherwise an LDS / wave-slot estimate).
int active_blocks(const Candidate& c, const DeviceInfo& dev);

}  // namespace lrt
