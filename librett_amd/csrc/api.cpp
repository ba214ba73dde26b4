// librett.h entry points, their C-ABI twins (librett_c.h) and plan storage.
//
// Reference counterpart: src/librett.cpp:45-398. Differences by design:
//  - handles come from one atomic fetch_add (the reference reads and then
//    increments, src/librett.cpp:128-129);
//  - plans live behind shared_ptr in a map guarded by a reader/writer lock,
//    and librettExecute holds the lock only for the lookup, not for the
//    launch (the reference holds its mutex through the launch, :375);
//  - HIP failures return LIBRETT_INTERNAL_ERROR instead of exiting
//    (src/GpuUtils.h:39-46);
//  - plans carry their kernel arguments by value, so creating one allocates
//    no device memory (the reference uploads tables in activate(),
//    src/plan.cpp:1648-1676).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/librett.h"
#include "../../include/librett_c.h"
#include "counters.h"
#include "kernels.h"
#include "lru_cache.h"
#include "planner.h"
#include "selector.h"

namespace lrt {
namespace {

struct PlanObj {
  int device = 0;
  hipStream_t stream = nullptr;
  Problem full;
  Candidate cand;
  Candidate fallback;  // no vector accesses: used for unaligned pointers
  int align = 1;       // pointer alignment the chosen candidate needs
  double measured_s = -1;
};

std::shared_mutex g_plans_mu;
std::unordered_map<unsigned int, std::shared_ptr<PlanObj>> g_plans;
std::atomic<unsigned int> g_next_handle{0};

std::mutex g_dev_mu;
std::unordered_map<int, DeviceInfo> g_devices;

struct ChoiceCacheValue {
  Candidate cand;
  double measured_s;
};
LruCache<std::string, ChoiceCacheValue> g_choice_cache(4096);

int env_int(const char* name, int def) {
  const char* v = std::getenv(name);
  if (!v || !*v) return def;
  return std::atoi(v);
}

// Validates the arguments the way librettPlanCheckInput does
// (src/librett.cpp:85-108), plus the 32-bit volume limit the kernels'
// index arithmetic assumes.
librettResult check_input(int rank, const int* dim, const int* perm, size_t sizeofType) {
  if (sizeofType != 4 && sizeofType != 8 && sizeofType != 16) return LIBRETT_INVALID_PARAMETER;
  if (rank < 1) return LIBRETT_INVALID_PARAMETER;
  if (!dim || !perm) return LIBRETT_INVALID_PARAMETER;
  long long vol = 1;
  for (int i = 0; i < rank; i++) {
    if (dim[i] < 1) return LIBRETT_INVALID_PARAMETER;
    vol *= dim[i];
    if (vol >= (1LL << 31)) return LIBRETT_INVALID_PARAMETER;
  }
  std::vector<char> seen(rank, 0);
  for (int i = 0; i < rank; i++) {
    if (perm[i] < 0 || perm[i] >= rank || seen[perm[i]]++) return LIBRETT_INVALID_PARAMETER;
  }
  return LIBRETT_SUCCESS;
}

bool device_info(DeviceInfo& out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  std::lock_guard<std::mutex> lock(g_dev_mu);
  auto it = g_devices.find(dev);
  if (it != g_devices.end()) {
    out = it->second;
    return true;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  DeviceInfo d;
  d.device_id = dev;
  d.cu_count = prop.multiProcessorCount;
  d.lds_per_block = (int)prop.sharedMemPerBlock;
  d.lds_per_cu = (int)std::max<size_t>(prop.sharedMemPerBlock, prop.maxSharedMemoryPerMultiProcessor);
  d.max_threads_per_block = prop.maxThreadsPerBlock;
  d.warp_size = prop.warpSize;
  d.max_waves_per_cu = std::max(1, prop.maxThreadsPerMultiProcessor / std::max(1, prop.warpSize));
  // l2CacheSize is the L2 of one XCD; MI355X has one XCD per 32 CUs.
  if (prop.l2CacheSize > 0) d.l2_bytes = (double)prop.l2CacheSize * std::max(1, d.cu_count / 32);
  d.has_gpu = 1;
  g_devices[dev] = d;
  out = d;
  return true;
}

DeviceInfo host_device(const librett_c_device* d) {
  DeviceInfo x;
  if (d) {
    x.cu_count = d->cu_count;
    x.lds_per_block = d->lds_per_block;
    x.lds_per_cu = d->lds_per_cu;
    x.max_threads_per_block = d->max_threads_per_block;
    x.max_waves_per_cu = d->max_waves_per_cu;
    x.hbm_bytes_per_s = d->hbm_bytes_per_s;
  }
  x.has_gpu = 0;
  return x;
}

std::string cache_key(const Problem& p, int device, bool measured) {
  std::string k;
  k.reserve(24 + p.rank * 8);
  k += measured ? 'M' : 'H';
  if (!measured) k += std::to_string(selector_generation()) + '.';  // heuristic choices depend on the selector
  k += std::to_string(device);
  k += ':';
  k += std::to_string(p.elem);
  for (int i = 0; i < p.rank; i++) {
    k += ',';
    k += std::to_string(p.dim[i]);
  }
  k += '/';
  for (int i = 0; i < p.rank; i++) {
    k += ',';
    k += std::to_string(p.perm[i]);
  }
  return k;
}

std::shared_ptr<PlanObj> make_plan_obj(const Problem& full, const Candidate& c, int device, hipStream_t stream) {
  auto p = std::make_shared<PlanObj>();
  p->device = device;
  p->stream = stream;
  p->full = full;
  p->cand = c;
  p->align = needed_alignment(c);
  p->fallback = unaligned_fallback(c);
  return p;
}

unsigned int store_plan(std::shared_ptr<PlanObj> p) {
  const unsigned int h = g_next_handle.fetch_add(1, std::memory_order_relaxed);
  std::unique_lock<std::shared_mutex> lock(g_plans_mu);
  g_plans[h] = std::move(p);
  return h;
}

std::shared_ptr<PlanObj> find_plan(unsigned int h) {
  std::shared_lock<std::shared_mutex> lock(g_plans_mu);
  auto it = g_plans.find(h);
  if (it == g_plans.end()) return nullptr;
  return it->second;
}

hipError_t run_plan(const PlanObj& p, const void* in, void* out) {
  const uintptr_t mis = ((uintptr_t)in | (uintptr_t)out) & (uintptr_t)(p.align - 1);
  return launch_candidate(mis ? p.fallback : p.cand, in, out, p.stream);
}

// Times one candidate: one warm-up launch, then `reps` launches each after an
// output memset, on the plan's stream. Returns the median in seconds, or a
// negative value on failure.
double time_candidate(const Candidate& c, hipStream_t stream, const void* in, void* out, int reps) {
  if (prepare_kernel(c) != 0) return -1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1;
  }
  std::vector<float> ts;
  bool ok = launch_candidate(c, in, out, stream) == hipSuccess;
  const size_t nbytes = (size_t)(c.bytes / 2);
  for (int r = 0; ok && r < reps; r++) {
    ok = hipMemsetAsync(out, 0xFF, nbytes, stream) == hipSuccess;
    ok = ok && hipEventRecord(e0, stream) == hipSuccess;
    ok = ok && launch_candidate(c, in, out, stream) == hipSuccess;
    ok = ok && hipEventRecord(e1, stream) == hipSuccess;
    ok = ok && hipEventSynchronize(e1) == hipSuccess;
    float ms = 0;
    ok = ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
    if (ok) ts.push_back(ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!ok || ts.empty()) {
    (void)hipGetLastError();
    return -1;
  }
  std::sort(ts.begin(), ts.end());
  return ts[ts.size() / 2] * 1e-3;
}

void fill_info(const Candidate& c, double measured, librett_c_plan_info* info) {
  std::memset(info, 0, sizeof(*info));
  info->method = c.method;
  info->elem = c.elem;
  info->tile_x = c.tile_x;
  info->tile_y = c.tile_y;
  info->vec_x = c.vec_x;
  info->vec_y = c.vec_y;
  info->word = c.word;
  info->copy_tx = c.copy_tx;
  info->nreg = c.nreg;
  info->num_split = c.split.numSplit;
  info->split_rank = c.split.splitRank;
  info->size_mm = c.split.sizeMm;
  info->size_mk = c.split.sizeMk;
  info->vol_mmk = c.split.volMmk;
  info->vol_mbar = c.split.volMbar;
  info->grid_x = c.grid_x;
  info->grid_y = c.grid_y;
  info->block = c.block;
  info->lds_bytes = c.lds_bytes;
  info->active_blocks = c.active_blocks_per_cu;
  info->predicted_s = c.model.predicted_s;
  info->measured_s = measured;
  info->read_lines = c.model.read_lines;
  info->write_lines = c.model.write_lines;
  info->partial_lines = c.model.partial_lines;
  info->bytes = c.bytes;
  info->feat[0] = c.model.read_tran;
  info->feat[1] = c.model.write_tran;
  info->feat[2] = c.model.items;
  info->feat[3] = c.model.concurrency;
  info->feat[4] = c.model.steps_per_block;
  info->feat[5] = c.model.t_hbm;
  info->feat[6] = c.model.t_latency;
  info->feat[7] = c.model.items > 0 ? (double)c.bytes / 2 / c.model.items : 0;
  std::snprintf(info->desc, sizeof(info->desc), "%s", c.describe().c_str());
}

// librettPlanMeasure's timing of every candidate (src/librett.cpp:224-343
// times each candidate once; here each gets a warm-up and the median of
// LIBRETT_MEASURE_REPS runs). Times what librettExecute would run on these
// pointers; a Trivial candidate is exclusive and ends the list. Negative
// entries are candidates that failed to run.
std::vector<double> measure_all(const Problem& full, const std::vector<Candidate>& cands, const DeviceInfo& dev,
                                hipStream_t stream, const void* in, void* out) {
  const int reps = std::max(1, env_int("LIBRETT_MEASURE_REPS", 3));
  std::vector<double> ts(cands.size(), -1.0);
  for (int i = 0; i < (int)cands.size(); i++) {
    auto probe = make_plan_obj(full, cands[i], dev.device_id, stream);
    const uintptr_t mis = ((uintptr_t)in | (uintptr_t)out) & (uintptr_t)(probe->align - 1);
    const Candidate& run = mis ? probe->fallback : probe->cand;
    ts[i] = time_candidate(run, stream, in, out, cands[i].method == kTrivial ? 1 : reps);
    if (cands[i].method == kTrivial) break;
  }
  return ts;
}

// Features of every candidate and the analytic model's choice.
void candidate_features(const Problem& full, const std::vector<Candidate>& cands, float* feats, int max) {
  const Problem red = reduce_ranks(full);
  for (int i = 0; i < (int)cands.size() && i < max; i++)
    selector_features(full, red, cands[i], feats + (size_t)i * kSelFeatures);
}

// Shared body of librettPlan / librettPlanMeasure / the forced-candidate plan.
// mode 0: model, 1: measure, 2: candidate `index`.
// LIBRETT_EXACT_COUNTS=1: before a model (or selector) choice, the line
// counts of Tiled and TiledCopy candidates whose model sampled only some of
// their tiles are replaced by the device counters' exact ones (the model
// counts lines the same way for these methods) and their times recomputed.
// Costs one counting launch per such candidate at plan time.
bool exact_counts_enabled() {
  static const bool on = env_int("LIBRETT_EXACT_COUNTS", 0) != 0;
  return on;
}

void refine_counts(std::vector<Candidate>& cands, const DeviceInfo& dev, hipStream_t stream) {
  for (Candidate& c : cands) {
    if ((c.method != kTiled && c.method != kTiledCopy) || c.model.items <= 24) continue;
    AccessCounts a;
    if (device_counts(c, stream, a) != hipSuccess) continue;
    c.model.read_lines = (double)a.v[kReadLines];
    c.model.write_lines = (double)a.v[kWriteLines];
    c.model.partial_lines = (double)a.v[kPartialLines];
    model_times(c, dev);
  }
}

librettResult plan_impl(unsigned int* handle, int rank, const int* dim, const int* perm, size_t sizeofType,
                        hipStream_t stream, int mode, const void* in, void* out, int index) {
  if (!handle) return LIBRETT_INVALID_PARAMETER;
  librettResult r = check_input(rank, dim, perm, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  if (mode == 1 && in == out) return LIBRETT_INVALID_PARAMETER;
  try {
    DeviceInfo dev;
    if (!device_info(dev)) return LIBRETT_INTERNAL_ERROR;
    const Problem full = canonicalize(rank, dim, perm, (int)sizeofType);
    const std::string key = cache_key(full, dev.device_id, mode == 1);
    Candidate chosen;
    double measured = -1;
    ChoiceCacheValue cached;
    if (mode != 2 && g_choice_cache.get(key, cached)) {
      chosen = cached.cand;
      measured = cached.measured_s;
    } else {
      std::vector<Candidate> cands = make_candidates(full, dev);
      if (cands.empty()) return LIBRETT_INTERNAL_ERROR;
      int best = -1;
      if (mode == 0) {
        if (exact_counts_enabled()) refine_counts(cands, dev, stream);
        // The learned selector when weights are loaded, else (or when it
        // declines the problem) the analytic model.
        best = selector_choose(cands, full, dev.device_id);
        if (best < 0) best = choose_by_model(cands);
      } else if (mode == 2) {
        if (index < 0 || index >= (int)cands.size()) return LIBRETT_INVALID_PARAMETER;
        best = index;
      } else {
        const std::vector<double> ts = measure_all(full, cands, dev, stream, in, out);
        double bt = 1e30;
        for (int i = 0; i < (int)ts.size(); i++)
          if (ts[i] >= 0 && ts[i] < bt) {
            bt = ts[i];
            best = i;
          }
        measured = bt;
      }
      if (best < 0) return LIBRETT_INTERNAL_ERROR;
      chosen = cands[best];
      if (mode != 2) g_choice_cache.put(key, ChoiceCacheValue{chosen, measured});
    }
    if (prepare_kernel(chosen) != 0) return LIBRETT_INTERNAL_ERROR;
    auto p = make_plan_obj(full, chosen, dev.device_id, stream);
    p->measured_s = measured;
    if (prepare_kernel(p->fallback) != 0) return LIBRETT_INTERNAL_ERROR;
    *handle = store_plan(std::move(p));
    return LIBRETT_SUCCESS;
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

}  // namespace
}  // namespace lrt

using namespace lrt;

// ---------------------------------------------------------------------------
// librett.h (C++ linkage, same mangled names as the reference)
// ---------------------------------------------------------------------------

// Optional (every entry point sets up what it needs on first use): takes the
// one-time set-up -- device properties, the planning threads, the learned
// selector's weights and inference buffers -- out of the first plan. The
// reference chooses its memory allocator here (src/librett.cpp:387-395);
// this library's allocator hook is librett_c_set_allocator.
void librettInitialize() {
  DeviceInfo dev;
  if (!device_info(dev)) return;
  planner_warm_up();
  kernels_warm_up();
  selector_prepare(dev.device_id);
}

void librettFinalize() {}

librettResult librettPlan(librettHandle* handle, int rank, int* dim, int* permutation, size_t sizeofType,
                          librett_gpuStream_t& stream) {
  return plan_impl(handle, rank, dim, permutation, sizeofType, stream, 0, nullptr, nullptr, -1);
}

librettResult librettPlanMeasure(librettHandle* handle, int rank, int* dim, int* permutation, size_t sizeofType,
                                 librett_gpuStream_t& stream, void* idata, void* odata) {
  return plan_impl(handle, rank, dim, permutation, sizeofType, stream, 1, idata, odata, -1);
}

librettResult librettDestroy(librettHandle handle) {
  std::unique_lock<std::shared_mutex> lock(g_plans_mu);
  auto it = g_plans.find(handle);
  if (it == g_plans.end()) return LIBRETT_INVALID_PLAN;
  g_plans.erase(it);
  return LIBRETT_SUCCESS;
}

librettResult librettExecute(librettHandle handle, void* idata, void* odata) {
  std::shared_ptr<PlanObj> p = find_plan(handle);
  if (!p) return LIBRETT_INVALID_PLAN;
  if (idata == odata) return LIBRETT_INVALID_PARAMETER;
  return run_plan(*p, idata, odata) == hipSuccess ? LIBRETT_SUCCESS : LIBRETT_INTERNAL_ERROR;
}

// ---------------------------------------------------------------------------
// librett_c.h
// ---------------------------------------------------------------------------

extern "C" {

void librett_c_initialize(void) { librettInitialize(); }

void librett_c_finalize(void) { librettFinalize(); }

int librett_c_plan(unsigned int* handle, int rank, const int* dim, const int* permutation, size_t sizeofType,
                   void* stream) {
  return plan_impl(handle, rank, dim, permutation, sizeofType, (hipStream_t)stream, 0, nullptr, nullptr, -1);
}

int librett_c_plan_measure(unsigned int* handle, int rank, const int* dim, const int* permutation,
                           size_t sizeofType, void* stream, void* idata, void* odata) {
  return plan_impl(handle, rank, dim, permutation, sizeofType, (hipStream_t)stream, 1, idata, odata, -1);
}

int librett_c_destroy(unsigned int handle) { return librettDestroy(handle); }

int librett_c_execute(unsigned int handle, const void* idata, void* odata) {
  return librettExecute(handle, const_cast<void*>(idata), odata);
}

int librett_c_plan_info_get(unsigned int handle, librett_c_plan_info* info) {
  if (!info) return LIBRETT_INVALID_PARAMETER;
  std::shared_ptr<PlanObj> p = find_plan(handle);
  if (!p) return LIBRETT_INVALID_PLAN;
  fill_info(p->cand, p->measured_s, info);
  return LIBRETT_SUCCESS;
}

int librett_c_candidates(int rank, const int* dim, const int* permutation, size_t sizeofType,
                         librett_c_plan_info* infos, int max, int* count) {
  librettResult r = check_input(rank, dim, permutation, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  try {
    DeviceInfo dev;
    if (!device_info(dev)) return LIBRETT_INTERNAL_ERROR;
    const Problem full = canonicalize(rank, dim, permutation, (int)sizeofType);
    std::vector<Candidate> cands = make_candidates(full, dev);
    if (count) *count = (int)cands.size();
    for (int i = 0; i < (int)cands.size() && i < max && infos; i++) fill_info(cands[i], -1, &infos[i]);
    return LIBRETT_SUCCESS;
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

int librett_c_plan_candidate(unsigned int* handle, int rank, const int* dim, const int* permutation,
                             size_t sizeofType, void* stream, int index) {
  return plan_impl(handle, rank, dim, permutation, sizeofType, (hipStream_t)stream, 2, nullptr, nullptr, index);
}

int librett_c_host_candidates(int rank, const int* dim, const int* permutation, size_t sizeofType,
                              const librett_c_device* d, librett_c_plan_info* infos, int max, int* count) {
  librettResult r = check_input(rank, dim, permutation, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  try {
    const DeviceInfo dev = host_device(d);
    const Problem full = canonicalize(rank, dim, permutation, (int)sizeofType);
    std::vector<Candidate> cands = make_candidates(full, dev);
    if (count) *count = (int)cands.size();
    for (int i = 0; i < (int)cands.size() && i < max && infos; i++) fill_info(cands[i], -1, &infos[i]);
    return LIBRETT_SUCCESS;
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

int librett_c_host_replay(int rank, const int* dim, const int* permutation, size_t sizeofType,
                          const librett_c_device* d, int index, const void* idata, void* odata) {
  librettResult r = check_input(rank, dim, permutation, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  if (!idata || !odata || idata == odata) return LIBRETT_INVALID_PARAMETER;
  try {
    const DeviceInfo dev = host_device(d);
    const Problem full = canonicalize(rank, dim, permutation, (int)sizeofType);
    std::vector<Candidate> cands = make_candidates(full, dev);
    if (index < 0 || index >= (int)cands.size()) return LIBRETT_INVALID_PARAMETER;
    // Guard bands around both buffers: a stray store shows up as a changed
    // output guard, a stray load as guard bytes in the output.
    const size_t n = (size_t)full.vol * sizeofType;
    const size_t g = 1 << 16;
    std::vector<unsigned char> in(n + 2 * g, 0xA5), out(n + 2 * g, 0x5A);
    std::memcpy(in.data() + g, idata, n);
    if (!replay_candidate(cands[index], in.data() + g, out.data() + g)) return LIBRETT_INTERNAL_ERROR;
    for (size_t i = 0; i < g; i++)
      if (out[i] != 0x5A || out[g + n + i] != 0x5A) return LIBRETT_UNDEFINED_ERROR;
    std::memcpy(odata, out.data() + g, n);
    return LIBRETT_SUCCESS;
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

int librett_c_measure_candidates(int rank, const int* dim, const int* permutation, size_t sizeofType, void* stream,
                                 const void* idata, void* odata, double* times, int max, int* count) {
  librettResult r = check_input(rank, dim, permutation, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  if (!idata || !odata || idata == odata) return LIBRETT_INVALID_PARAMETER;
  try {
    DeviceInfo dev;
    if (!device_info(dev)) return LIBRETT_INTERNAL_ERROR;
    const Problem full = canonicalize(rank, dim, permutation, (int)sizeofType);
    const std::vector<Candidate> cands = make_candidates(full, dev);
    const std::vector<double> ts = measure_all(full, cands, dev, (hipStream_t)stream, idata, odata);
    if (count) *count = (int)cands.size();
    for (int i = 0; i < (int)ts.size() && i < max && times; i++) times[i] = ts[i];
    return LIBRETT_SUCCESS;
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

int librett_c_selector_features(int rank, const int* dim, const int* permutation, size_t sizeofType,
                                const librett_c_device* d, float* feats, int max, int* count, int* model_choice) {
  librettResult r = check_input(rank, dim, permutation, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  try {
    DeviceInfo dev;
    if (d)
      dev = host_device(d);
    else if (!device_info(dev))
      return LIBRETT_INTERNAL_ERROR;
    const Problem full = canonicalize(rank, dim, permutation, (int)sizeofType);
    const std::vector<Candidate> cands = make_candidates(full, dev);
    if (count) *count = (int)cands.size();
    if (model_choice) *model_choice = choose_by_model(cands);
    if (feats) candidate_features(full, cands, feats, max);
    return LIBRETT_SUCCESS;
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

int librett_c_selector_choose(int rank, const int* dim, const int* permutation, size_t sizeofType, int* index) {
  librettResult r = check_input(rank, dim, permutation, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  if (!index) return LIBRETT_INVALID_PARAMETER;
  try {
    DeviceInfo dev;
    if (!device_info(dev)) return LIBRETT_INTERNAL_ERROR;
    const Problem full = canonicalize(rank, dim, permutation, (int)sizeofType);
    const std::vector<Candidate> cands = make_candidates(full, dev);
    *index = selector_choose(cands, full, dev.device_id);
    return LIBRETT_SUCCESS;
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

int librett_c_reduce_ranks(int rank, const int* dim, const int* permutation, int* red_rank, int* red_dim,
                           int* red_perm) {
  librettResult r = check_input(rank, dim, permutation, 4);
  if (r != LIBRETT_SUCCESS) return r;
  const Problem full = canonicalize(rank, dim, permutation, 4);
  const Problem red = reduce_ranks(full);
  if (red_rank) *red_rank = red.rank;
  for (int i = 0; i < red.rank; i++) {
    if (red_dim) red_dim[i] = red.dim[i];
    if (red_perm) red_perm[i] = red.perm[i];
  }
  return LIBRETT_SUCCESS;
}

int librett_c_device_copy(void* dst, const void* src, size_t bytes, void* stream, int words_per_thread,
                          int nontemporal, int grid) {
  const hipError_t e =
      launch_device_copy(dst, src, bytes, (hipStream_t)stream, words_per_thread, nontemporal, grid);
  if (e == hipErrorInvalidValue) {
    (void)hipGetLastError();
    return LIBRETT_INVALID_PARAMETER;
  }
  return e == hipSuccess ? LIBRETT_SUCCESS : LIBRETT_INTERNAL_ERROR;
}

void librett_c_set_model_constants(const double* values, int count) { set_model_constants(values, count); }

const char* librett_c_version(void) { return "librett-mi355x 0.1 (gfx950)"; }

static void copy_counts(const AccessCounts& a, librett_c_counters* out) {
  out->ld_req = a.v[kLdReq];
  out->st_req = a.v[kStReq];
  out->ld_tran = a.v[kLdTran];
  out->st_tran = a.v[kStTran];
  out->ld_bytes = a.v[kLdBytes];
  out->st_bytes = a.v[kStBytes];
  out->read_lines = a.v[kReadLines];
  out->write_lines = a.v[kWriteLines];
  out->partial_lines = a.v[kPartialLines];
  out->items = a.v[kItems];
}

static int counts_of(const Candidate& c, hipStream_t stream, librett_c_counters* out) {
  if (c.method == kTrivial) return LIBRETT_INVALID_PARAMETER;
  AccessCounts a;
  const hipError_t e = device_counts(c, stream, a);
  if (e == hipErrorInvalidValue) return LIBRETT_INVALID_PARAMETER;
  if (e != hipSuccess) return LIBRETT_INTERNAL_ERROR;
  copy_counts(a, out);
  return LIBRETT_SUCCESS;
}

int librett_c_plan_counters(unsigned int handle, librett_c_counters* out) {
  if (!out) return LIBRETT_INVALID_PARAMETER;
  std::shared_ptr<PlanObj> p = find_plan(handle);
  if (!p) return LIBRETT_INVALID_PLAN;
  try {
    return counts_of(p->cand, p->stream, out);
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

int librett_c_candidate_counters(int rank, const int* dim, const int* permutation, size_t sizeofType, void* stream,
                                 int index, librett_c_counters* out) {
  librettResult r = check_input(rank, dim, permutation, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  if (!out) return LIBRETT_INVALID_PARAMETER;
  try {
    DeviceInfo dev;
    if (!device_info(dev)) return LIBRETT_INTERNAL_ERROR;
    const Problem full = canonicalize(rank, dim, permutation, (int)sizeofType);
    const std::vector<Candidate> cands = make_candidates(full, dev);
    if (index < 0 || index >= (int)cands.size()) return LIBRETT_INVALID_PARAMETER;
    return counts_of(cands[index], (hipStream_t)stream, out);
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

int librett_c_host_candidate_counters(int rank, const int* dim, const int* permutation, size_t sizeofType,
                                      const librett_c_device* d, int index, librett_c_counters* out) {
  librettResult r = check_input(rank, dim, permutation, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  if (!out) return LIBRETT_INVALID_PARAMETER;
  try {
    const DeviceInfo dev = host_device(d);
    const Problem full = canonicalize(rank, dim, permutation, (int)sizeofType);
    const std::vector<Candidate> cands = make_candidates(full, dev);
    if (index < 0 || index >= (int)cands.size() || cands[index].method == kTrivial) return LIBRETT_INVALID_PARAMETER;
    AccessCounts a;
    if (!host_counts(cands[index], a)) return LIBRETT_INVALID_PARAMETER;
    copy_counts(a, out);
    return LIBRETT_SUCCESS;
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

int librett_c_host_access_trace(int rank, const int* dim, const int* permutation, size_t sizeofType,
                                const librett_c_device* d, int index, int source, long long* records,
                                long long max, long long* count) {
  librettResult r = check_input(rank, dim, permutation, sizeofType);
  if (r != LIBRETT_SUCCESS) return r;
  if (!count || (source != 0 && source != 1)) return LIBRETT_INVALID_PARAMETER;
  try {
    const DeviceInfo dev = host_device(d);
    const Problem full = canonicalize(rank, dim, permutation, (int)sizeofType);
    const std::vector<Candidate> cands = make_candidates(full, dev);
    if (index < 0 || index >= (int)cands.size() || cands[index].method == kTrivial) return LIBRETT_INVALID_PARAMETER;
    std::vector<AccessRecord> recs;
    const bool ok = source == 0 ? trace_enumeration(cands[index], recs) : trace_kernel(cands[index], recs);
    if (!ok) return LIBRETT_INTERNAL_ERROR;
    *count = (long long)recs.size();
    for (long long i = 0; i < (long long)recs.size() && i < max && records; i++) {
      long long* q = records + 6 * i;
      q[0] = recs[i].item;
      q[1] = recs[i].store;
      q[2] = recs[i].inst;
      q[3] = recs[i].thread;
      q[4] = recs[i].offset;
      q[5] = recs[i].bytes;
    }
    return LIBRETT_SUCCESS;
  } catch (...) {
    return LIBRETT_INTERNAL_ERROR;
  }
}

}  // extern "C"
