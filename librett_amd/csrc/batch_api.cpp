// Batched engine: host side (planning, autotuning, descriptor arena, HIP
// graph) and its C-ABI entry points (librett_c.h). The reference has no
// batched mode; its per-call path is librettExecute (src/librett.cpp:372-385),
// which is the baseline this engine is measured against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "devmem.h"
#include "../../include/librett_c.h"
#include "batch.h"
#include "kernels.h"
#include "lru_cache.h"
#include "planner.h"

namespace lrt {
namespace {

// Internal failures name their cause on stderr when LIBRETT_DEBUG is set.
int fail(const char* why, int code = LIBRETT_C_INTERNAL_ERROR) {
  static const bool on = std::getenv("LIBRETT_DEBUG") != nullptr;
  if (on) std::fprintf(stderr, "librett batch: %s\n", why);
  return code;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int check_problem(int rank, const int* dim, const int* perm, size_t elem) {
  if (elem != 4 && elem != 8 && elem != 16) return LIBRETT_C_INVALID_PARAMETER;
  if (rank < 1 || !dim || !perm) return LIBRETT_C_INVALID_PARAMETER;
  long long vol = 1;
  std::vector<char> seen(rank, 0);
  for (int i = 0; i < rank; i++) {
    if (dim[i] < 1) return LIBRETT_C_INVALID_PARAMETER;
    vol *= dim[i];
    if (vol >= (1LL << 31)) return LIBRETT_C_INVALID_PARAMETER;
    if (perm[i] < 0 || perm[i] >= rank || seen[perm[i]]++) return LIBRETT_C_INVALID_PARAMETER;
  }
  return LIBRETT_C_SUCCESS;
}

// A flat copy of `bytes` bytes in words of W bytes as TiledCopy
// descriptors: 64-word rows, plus a one-row descriptor for the tail. Used for
// problems that reduce to a plain copy (the reference's Trivial method).
void copy_descs(long long bytes, int W, std::vector<BatchDesc>& out, std::vector<long long>& byte_offsets) {
  const long long n = bytes / W;
  const long long rows = n / 64, tail = n % 64;
  auto make = [&](int X, long long Y, long long off_words) {
    Candidate c;
    c.method = kTiledCopy;
    c.word = W;
    c.elem = W;
    c.copy_tx = X <= 16 ? 16 : 64;
    CopyArgs& a = c.copy;
    a.X = X;
    a.Y = (int)Y;
    a.in_ys = X;
    a.out_ys = X;
    const int ty = (256 / c.copy_tx) * 4;
    a.ntx = (X + c.copy_tx - 1) / c.copy_tx;
    const int nty = (int)((Y + ty - 1) / ty);
    a.ntiles = a.ntx * nty;
    a.ntx_div = make_fastdiv((uint32_t)a.ntx);
    a.ntiles_div = make_fastdiv((uint32_t)a.ntiles);
    a.items = a.ntiles;
    a.outer.n = 0;
    for (int r = 0; r < kMaxRank; r++) {
      a.outer.div[r] = make_fastdiv(1);
      a.outer.ct_in[r] = a.outer.ct_out[r] = 0;
    }
    BatchDesc d;
    if (batch_describe(c, d)) {
      out.push_back(d);
      byte_offsets.push_back(off_words * W);
    }
  };
  if (rows > 0) make(64, rows, 0);
  if (tail > 0) make((int)tail, 1, rows * 64);
}

struct BatchObj {
  int device = 0;
  hipStream_t stream = nullptr;
  int nprob = 0;
  std::vector<Problem> probs;
  std::vector<Candidate> chosen;
  // Descriptors: one or two per problem (a flat copy may need a tail).
  std::vector<BatchDesc> descs, fallback;
  std::vector<int> owner;               // problem of each descriptor
  std::vector<long long> byte_offset;   // offset of the descriptor's data in its problem
  BatchLayout layout;                   // descriptors grouped by kernel family
  long long total = 0;                  // units of all families
  int grid[kBatchFamilies] = {}; // workgroups of each family's launch
  int grid_max = 0;
  long long unit_bytes = kBatchUnitBytes;
  BatchDesc* d_descs = nullptr;
  int* d_start = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool buffers_set = false;
  double plan_s = 0, tune_s = 0;
  int tuned = 0;

  ~BatchObj() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (d_descs) (void)dev_free(d_descs);
    if (d_start) (void)dev_free(d_start);
  }
};

std::mutex g_batch_mu;
std::unordered_map<unsigned int, std::shared_ptr<BatchObj>> g_batches;
std::atomic<unsigned int> g_next_batch{1};

struct TuneValue {
  Candidate cand;
  double seconds;
};
LruCache<std::string, TuneValue> g_tune_cache(1 << 16);

std::string problem_key(const Problem& p) {
  std::string k = std::to_string(p.elem) + ":";
  for (int i = 0; i < p.rank; i++) k += std::to_string(p.dim[i]) + ",";
  k += "/";
  for (int i = 0; i < p.rank; i++) k += std::to_string(p.perm[i]) + ",";
  return k;
}


// Times candidate c of problem p inside the persistent kernel, on enough
// clones of the problem to fill the device (scratch buffers `a`, `b`).
double time_in_batch(const Candidate& c, const Problem& p, void* a, void* b, size_t scratch, hipStream_t stream,
                     BatchDesc* d_desc, int* d_start, int max_clones, const int* grids, long long unit_bytes) {
  BatchDesc d;
  if (!batch_describe(c, d)) return -1;
  batch_regroup(d, unit_bytes);
  const int fam = batch_family(d.variant);
  const int grid = grids[fam];
  const size_t bytes = (size_t)p.vol * p.elem;
  const size_t span = (bytes + 255) / 256 * 256;
  if (span > scratch) return -1;  // every clone must lie inside the scratch buffers
  int clones = (int)std::min<size_t>(max_clones, std::max<size_t>(1, scratch / span));
  clones = std::max(1, std::min(clones, (int)((64ull << 20) / std::max<size_t>(1, bytes)) + 1));
  std::vector<BatchDesc> ds(clones, d);
  std::vector<int> st(clones + 1);
  for (int k = 0; k < clones; k++) {
    ds[k].in = (const char*)a + k * span;
    ds[k].out = (char*)b + k * span;
    st[k] = k * d.units;
  }
  st[clones] = clones * d.units;
  if (hipMemcpyAsync(d_desc, ds.data(), clones * sizeof(BatchDesc), hipMemcpyHostToDevice, stream) != hipSuccess)
    return -1;
  if (hipMemcpyAsync(d_start, st.data(), (clones + 1) * sizeof(int), hipMemcpyHostToDevice, stream) != hipSuccess)
    return -1;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return -1;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return -1;
  }
  bool ok = launch_batch(fam, d_desc, d_start, clones, st[clones], grid, stream) == hipSuccess;
  float best = 1e30f;
  for (int r = 0; ok && r < 2; r++) {
    ok = hipEventRecord(e0, stream) == hipSuccess &&
         launch_batch(fam, d_desc, d_start, clones, st[clones], grid, stream) == hipSuccess &&
         hipEventRecord(e1, stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
    float ms = 0;
    if (ok && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) best = std::min(best, ms);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!ok) {
    (void)hipGetLastError();
    return -1;
  }
  return best * 1e-3 / clones;
}

int build_batch(BatchObj& B, const DeviceInfo& dev, int tune) {
  const double t0 = now_s();
  DeviceInfo bdev = dev;
  bdev.batch_mode = 1;
  B.chosen.resize(B.nprob);
  std::vector<std::vector<Candidate>> ranked(B.nprob);
  for (int i = 0; i < B.nprob; i++) {
    ranked[i] = batch_candidates(B.probs[i], bdev);
    if (ranked[i].empty()) return fail("no batch-mode candidate");
    B.chosen[i] = ranked[i][0];
  }
  B.plan_s = now_s() - t0;

  B.grid_max = 0;
  for (int f = 0; f < kBatchFamilies; f++) {
    const int occ = std::max(1, dev.has_gpu ? batch_occupancy(f) : 4);
    B.grid[f] = dev.cu_count * occ;
    B.grid_max = std::max(B.grid_max, B.grid[f]);
  }
  long long total_bytes = 0;
  for (const Problem& p : B.probs) total_bytes += 2LL * p.vol * p.elem;
  B.unit_bytes = batch_unit_bytes(total_bytes, B.grid_max);

  if (tune && dev.has_gpu) {
    // Autotuner: time the top candidates of each distinct problem inside the
    // persistent kernel on clones of the problem, keep the fastest, and
    // remember it across batches.
    const double t1 = now_s();
    // Scratch for the clones: at least 256 MiB, and room for the largest
    // problem up to 4 GiB (larger problems keep the model's choice).
    size_t largest = 0;
    for (const Problem& p : B.probs) largest = std::max(largest, ((size_t)p.vol * p.elem + 255) / 256 * 256);
    const size_t scratch = std::max<size_t>(256ull << 20, std::min<size_t>(largest, 4ull << 30));
    void *a = nullptr, *b = nullptr;
    BatchDesc* d_desc = nullptr;
    int* d_start = nullptr;
    const int max_clones = 4096;
    bool ok = dev_malloc(&a, scratch) == hipSuccess && dev_malloc(&b, scratch) == hipSuccess &&
              dev_malloc((void**)&d_desc, max_clones * sizeof(BatchDesc)) == hipSuccess &&
              dev_malloc((void**)&d_start, (max_clones + 1) * sizeof(int)) == hipSuccess;
    if (ok) ok = hipMemsetAsync(a, 0, scratch, B.stream) == hipSuccess;
    const int topk = std::max(1, std::min(8, tune));
    for (int i = 0; ok && i < B.nprob; i++) {
      if (B.chosen[i].method == kTrivial) continue;
      const std::string key = problem_key(B.probs[i]);
      TuneValue tv;
      if (g_tune_cache.get(key, tv)) {
        B.chosen[i] = tv.cand;
        continue;
      }
      double best = 1e30;
      int bi = 0;
      for (int k = 0; k < (int)ranked[i].size() && k < topk; k++) {
        const double t = time_in_batch(ranked[i][k], B.probs[i], a, b, scratch, B.stream, d_desc, d_start,
                                       max_clones, B.grid, B.unit_bytes);
        if (t > 0 && t < best) {
          best = t;
          bi = k;
        }
      }
      B.chosen[i] = ranked[i][bi];
      B.tuned++;
      g_tune_cache.put(key, TuneValue{B.chosen[i], best});
    }
    (void)hipStreamSynchronize(B.stream);
    if (a) (void)dev_free(a);
    if (b) (void)dev_free(b);
    if (d_desc) (void)dev_free(d_desc);
    if (d_start) (void)dev_free(d_start);
    if (!ok) return fail("autotuner HIP call failed");
    B.tune_s = now_s() - t1;
  }

  // Descriptors and the unit index space.
  B.descs.clear();
  B.fallback.clear();
  B.owner.clear();
  B.byte_offset.clear();
  for (int i = 0; i < B.nprob; i++) {
    std::vector<BatchDesc> ds, fs;
    std::vector<long long> offs;
    if (!describe_problem(B.probs[i], B.chosen[i], ds, fs, offs)) return fail("no descriptor for the chosen plan");
    for (size_t k = 0; k < ds.size(); k++) {
      batch_regroup(ds[k], B.unit_bytes);
      batch_regroup(fs[k], B.unit_bytes);
      B.descs.push_back(ds[k]);
      B.fallback.push_back(fs[k]);
      B.owner.push_back(i);
      B.byte_offset.push_back(offs[k]);
    }
  }
  // Group by kernel family; the fallback descriptors, owners and offsets
  // follow the same order (a fallback keeps its descriptor's family).
  std::vector<int> order;
  batch_layout(B.descs, B.layout, &order);
  std::vector<BatchDesc> fb(order.size());
  std::vector<int> own(order.size());
  std::vector<long long> off(order.size());
  for (size_t k = 0; k < order.size(); k++) {
    fb[k] = B.fallback[order[k]];
    own[k] = B.owner[order[k]];
    off[k] = B.byte_offset[order[k]];
  }
  B.fallback.swap(fb);
  B.owner.swap(own);
  B.byte_offset.swap(off);
  B.total = 0;
  for (int f = 0; f < kBatchFamilies; f++) {
    if (B.layout.total[f] >= (1LL << 31)) return LIBRETT_C_INVALID_PARAMETER;
    B.total += B.layout.total[f];
  }
  return LIBRETT_C_SUCCESS;
}

std::shared_ptr<BatchObj> find_batch(unsigned int h) {
  std::lock_guard<std::mutex> lock(g_batch_mu);
  auto it = g_batches.find(h);
  return it == g_batches.end() ? nullptr : it->second;
}

bool device_info_batch(DeviceInfo& d) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  d.device_id = dev;
  d.cu_count = prop.multiProcessorCount;
  d.lds_per_block = (int)prop.sharedMemPerBlock;
  d.lds_per_cu = (int)std::max<size_t>(prop.sharedMemPerBlock, prop.maxSharedMemoryPerMultiProcessor);
  d.max_threads_per_block = prop.maxThreadsPerBlock;
  d.warp_size = prop.warpSize;
  d.max_waves_per_cu = std::max(1, prop.maxThreadsPerMultiProcessor / std::max(1, prop.warpSize));
  if (prop.l2CacheSize > 0) d.l2_bytes = (double)prop.l2CacheSize * std::max(1, d.cu_count / 32);
  d.has_gpu = 1;
  return true;
}

int read_problems(int count, const int* ranks, const int* dims, const int* perms, const size_t* elems,
                  std::vector<Problem>& out) {
  if (count < 1 || !ranks || !dims || !perms || !elems) return LIBRETT_C_INVALID_PARAMETER;
  out.clear();
  long long off = 0;
  for (int i = 0; i < count; i++) {
    const int r = check_problem(ranks[i], dims + off, perms + off, elems[i]);
    if (r != LIBRETT_C_SUCCESS) return r;
    out.push_back(canonicalize(ranks[i], dims + off, perms + off, (int)elems[i]));
    off += ranks[i];
  }
  return LIBRETT_C_SUCCESS;
}

}  // namespace

std::vector<Candidate> batch_candidates(const Problem& p, const DeviceInfo& dev) {
  DeviceInfo bdev = dev;
  bdev.batch_mode = 1;
  std::vector<Candidate> cands = make_candidates(p, bdev);
  std::vector<Candidate> ok;
  for (auto& c : cands)
    if (c.method == kTrivial || batch_variant(c) >= 0) ok.push_back(c);
  std::stable_sort(ok.begin(), ok.end(), [](const Candidate& a, const Candidate& b) {
    return a.model.t_stream < b.model.t_stream;
  });
  return ok;
}

int batch_alignment(const Problem& p, const Candidate& c) {
  if (c.method == kTrivial) return p.elem;
  if (c.method == kTiled) return c.elem * std::max(c.vec_x, c.vec_y);
  if (c.method == kTiledCopy) return c.word;
  return c.elem;
}

bool describe_problem(const Problem& p, const Candidate& c, std::vector<BatchDesc>& descs,
                      std::vector<BatchDesc>& fallback, std::vector<long long>& offsets) {
  descs.clear();
  fallback.clear();
  offsets.clear();
  if (c.method == kTrivial) {
    // Element-sized words: always aligned, so the fallback is the same.
    copy_descs((long long)p.vol * p.elem, p.elem, descs, offsets);
    fallback = descs;
    return !descs.empty();
  }
  BatchDesc d, f;
  if (!batch_describe(c, d)) return false;
  Candidate fc = c;
  if (c.method == kTiled) fc.vec_x = fc.vec_y = 1;
  if (c.method == kTiledCopy && c.word > c.elem) {
    // Element-sized words for unaligned buffers.
    const int k = c.word / c.elem;
    CopyArgs& a = fc.copy;
    a.X *= k;
    a.in_ys *= k;
    a.out_ys *= k;
    for (int r = 0; r < a.outer.n; r++) {
      a.outer.ct_in[r] *= k;
      a.outer.ct_out[r] *= k;
    }
    fc.word = c.elem;
    fc.copy_tx = 64;
    const int ty = (256 / fc.copy_tx) * 4;
    const long long vol_outer = c.copy.items / std::max(1, c.copy.ntiles);
    a.ntx = (a.X + fc.copy_tx - 1) / fc.copy_tx;
    a.ntiles = a.ntx * ((a.Y + ty - 1) / ty);
    a.ntx_div = make_fastdiv((uint32_t)a.ntx);
    a.ntiles_div = make_fastdiv((uint32_t)a.ntiles);
    a.items = (int)(a.ntiles * vol_outer);
  }
  if (!batch_describe(fc, f)) f = d;
  descs.push_back(d);
  fallback.push_back(f);
  offsets.push_back(0);
  return true;
}

}  // namespace lrt

using namespace lrt;

extern "C" {

int librett_c_batch_plan(unsigned int* handle, int count, const int* ranks, const int* dims, const int* perms,
                         const size_t* elems, void* stream, int tune) {
  if (!handle) return LIBRETT_C_INVALID_PARAMETER;
  try {
    auto B = std::make_shared<BatchObj>();
    int r = read_problems(count, ranks, dims, perms, elems, B->probs);
    if (r != LIBRETT_C_SUCCESS) return r;
    B->nprob = count;
    B->stream = (hipStream_t)stream;
    DeviceInfo dev;
    if (!device_info_batch(dev)) return fail("device query failed");
    B->device = dev.device_id;
    r = build_batch(*B, dev, tune);
    if (r != LIBRETT_C_SUCCESS) return r;
    const size_t nd = B->descs.size();
    if (dev_malloc((void**)&B->d_descs, nd * sizeof(BatchDesc)) != hipSuccess ||
        dev_malloc((void**)&B->d_start, B->layout.start.size() * sizeof(int)) != hipSuccess) {
      (void)hipGetLastError();
      return fail("descriptor allocation failed");
    }
    if (hipMemcpy(B->d_start, B->layout.start.data(), B->layout.start.size() * sizeof(int), hipMemcpyHostToDevice) !=
        hipSuccess)
      return fail("descriptor upload failed");
    const unsigned int h = g_next_batch.fetch_add(1);
    {
      std::lock_guard<std::mutex> lock(g_batch_mu);
      g_batches[h] = B;
    }
    *handle = h;
    return LIBRETT_C_SUCCESS;
  } catch (const std::exception& e) {
    return fail(e.what());
  } catch (...) {
    return fail("unknown exception");
  }
}

int librett_c_batch_set_buffers(unsigned int handle, const void* const* idata, void* const* odata) {
  auto B = find_batch(handle);
  if (!B) return LIBRETT_C_INVALID_PLAN;
  if (!idata || !odata) return LIBRETT_C_INVALID_PARAMETER;
  std::vector<BatchDesc> up(B->descs.size());
  for (size_t k = 0; k < B->descs.size(); k++) {
    const int i = B->owner[k];
    if (!idata[i] || !odata[i] || idata[i] == odata[i]) return LIBRETT_C_INVALID_PARAMETER;
    const char* in = (const char*)idata[i] + B->byte_offset[k];
    char* out = (char*)odata[i] + B->byte_offset[k];
    const int align = batch_alignment(B->probs[i], B->chosen[i]);
    const bool mis = (((uintptr_t)in | (uintptr_t)out) & (uintptr_t)(align - 1)) != 0;
    up[k] = mis ? B->fallback[k] : B->descs[k];
    up[k].in = in;
    up[k].out = out;
  }
  // A fallback descriptor may have another unit count: rebuild the index
  // space and re-capture the graph when it changes. The descriptors are
  // already in family order, and a fallback keeps its family, so the
  // layout keeps that order.
  BatchLayout L;
  std::vector<BatchDesc> tmp = up;
  batch_layout(tmp, L, nullptr);
  long long total = 0;
  for (int f = 0; f < kBatchFamilies; f++) {
    if (L.total[f] >= (1LL << 31)) return LIBRETT_C_INVALID_PARAMETER;
    total += L.total[f];
  }
  if (L.start != B->layout.start) {
    B->layout = L;
    B->total = total;
    if (hipMemcpyAsync(B->d_start, L.start.data(), L.start.size() * sizeof(int), hipMemcpyHostToDevice,
                       B->stream) != hipSuccess)
      return LIBRETT_C_INTERNAL_ERROR;
    if (B->exec) {
      (void)hipGraphExecDestroy(B->exec);
      B->exec = nullptr;
    }
    if (B->graph) {
      (void)hipGraphDestroy(B->graph);
      B->graph = nullptr;
    }
  }
  if (hipMemcpyAsync(B->d_descs, up.data(), up.size() * sizeof(BatchDesc), hipMemcpyHostToDevice, B->stream) !=
          hipSuccess ||
      hipStreamSynchronize(B->stream) != hipSuccess) {
    (void)hipGetLastError();
    return LIBRETT_C_INTERNAL_ERROR;
  }
  B->buffers_set = true;
  return LIBRETT_C_SUCCESS;
}

int librett_c_batch_execute(unsigned int handle, int use_graph) {
  auto B = find_batch(handle);
  if (!B) return LIBRETT_C_INVALID_PLAN;
  if (!B->buffers_set) return LIBRETT_C_INVALID_PARAMETER;
  if (!use_graph) {
    return launch_batch_all(B->d_descs, B->d_start, B->layout, B->grid, B->stream) == hipSuccess
               ? LIBRETT_C_SUCCESS
               : LIBRETT_C_INTERNAL_ERROR;
  }
  if (!B->exec) {
    if (hipStreamBeginCapture(B->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      return LIBRETT_C_INTERNAL_ERROR;
    }
    const hipError_t le = launch_batch_all(B->d_descs, B->d_start, B->layout, B->grid, B->stream);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(B->stream, &g);
    if (le != hipSuccess || ce != hipSuccess || !g) {
      (void)hipGetLastError();
      if (g) (void)hipGraphDestroy(g);
      return LIBRETT_C_INTERNAL_ERROR;
    }
    hipGraphExec_t ex = nullptr;
    if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGetLastError();
      (void)hipGraphDestroy(g);
      return LIBRETT_C_INTERNAL_ERROR;
    }
    B->graph = g;
    B->exec = ex;
  }
  return hipGraphLaunch(B->exec, B->stream) == hipSuccess ? LIBRETT_C_SUCCESS : LIBRETT_C_INTERNAL_ERROR;
}

int librett_c_batch_destroy(unsigned int handle) {
  std::lock_guard<std::mutex> lock(g_batch_mu);
  auto it = g_batches.find(handle);
  if (it == g_batches.end()) return LIBRETT_C_INVALID_PLAN;
  g_batches.erase(it);
  return LIBRETT_C_SUCCESS;
}

int librett_c_batch_info_get(unsigned int handle, librett_c_batch_info* info) {
  auto B = find_batch(handle);
  if (!B) return LIBRETT_C_INVALID_PLAN;
  if (!info) return LIBRETT_C_INVALID_PARAMETER;
  std::memset(info, 0, sizeof(*info));
  info->count = B->nprob;
  info->descriptors = (int)B->descs.size();
  info->total_units = (int)std::min<long long>(B->total, 0x7fffffff);
  info->grid = B->grid_max;
  info->tuned = B->tuned;
  info->plan_seconds = B->plan_s;
  info->tune_seconds = B->tune_s;
  for (const auto& c : B->chosen)
    if (c.method >= 0 && c.method < 6) info->method_count[c.method]++;
  for (int i = 0; i < B->nprob; i++) info->bytes += B->chosen[i].bytes;
  return LIBRETT_C_SUCCESS;
}

int librett_c_batch_host_replay(int count, const int* ranks, const int* dims, const int* perms, const size_t* elems,
                                const librett_c_device* d, const void* const* idata, void* const* odata, int grid) {
  try {
    BatchObj B;
    int r = read_problems(count, ranks, dims, perms, elems, B.probs);
    if (r != LIBRETT_C_SUCCESS) return r;
    B.nprob = count;
    DeviceInfo dev;
    if (d) {
      dev.cu_count = d->cu_count;
      dev.lds_per_block = d->lds_per_block;
      dev.lds_per_cu = d->lds_per_cu;
      dev.max_threads_per_block = d->max_threads_per_block;
      dev.max_waves_per_cu = d->max_waves_per_cu;
      dev.hbm_bytes_per_s = d->hbm_bytes_per_s;
    }
    dev.has_gpu = 0;
    r = build_batch(B, dev, 0);
    if (r != LIBRETT_C_SUCCESS) return r;
    // Every problem runs between guard bands; a stray store changes a
    // guard, a stray load brings guard bytes into an output.
    const size_t g = 4096;
    std::vector<std::vector<unsigned char>> gin(count), gout(count);
    for (int i = 0; i < count; i++) {
      const size_t n = (size_t)B.probs[i].vol * B.probs[i].elem;
      gin[i].assign(n + 2 * g, 0xA5);
      gout[i].assign(n + 2 * g, 0x5A);
      std::memcpy(gin[i].data() + g, idata[i], n);
    }
    std::vector<BatchDesc> ds = B.descs;
    for (size_t k = 0; k < ds.size(); k++) {
      const int i = B.owner[k];
      ds[k].in = gin[i].data() + g + B.byte_offset[k];
      ds[k].out = gout[i].data() + g + B.byte_offset[k];
    }
    if (!replay_batch_all(ds.data(), B.layout, grid > 0 ? grid : B.grid_max)) return LIBRETT_C_UNDEFINED_ERROR;
    for (int i = 0; i < count; i++) {
      const size_t n = (size_t)B.probs[i].vol * B.probs[i].elem;
      for (size_t j = 0; j < g; j++)
        if (gout[i][j] != 0x5A || gout[i][g + n + j] != 0x5A) return LIBRETT_C_UNDEFINED_ERROR;
      std::memcpy(odata[i], gout[i].data() + g, n);
    }
    return LIBRETT_C_SUCCESS;
  } catch (...) {
    return LIBRETT_C_INTERNAL_ERROR;
  }
}

int librett_c_batch_host_candidate(int rank, const int* dim, const int* permutation, size_t sizeofType,
                                   const librett_c_device* d, int index, int fallback, long long unit_bytes,
                                   int grid, const void* idata, void* odata, int* count) {
  try {
    std::vector<Problem> probs;
    const size_t elems[1] = {sizeofType};
    int r = read_problems(1, &rank, dim, permutation, elems, probs);
    if (r != LIBRETT_C_SUCCESS) return r;
    DeviceInfo dev;
    if (d) {
      dev.cu_count = d->cu_count;
      dev.lds_per_block = d->lds_per_block;
      dev.lds_per_cu = d->lds_per_cu;
      dev.max_threads_per_block = d->max_threads_per_block;
      dev.max_waves_per_cu = d->max_waves_per_cu;
      dev.hbm_bytes_per_s = d->hbm_bytes_per_s;
    }
    dev.has_gpu = 0;
    const Problem& p = probs[0];
    const std::vector<Candidate> ranked = batch_candidates(p, dev);
    if (count) *count = (int)ranked.size();
    if (!idata && !odata) return LIBRETT_C_SUCCESS;
    if (!idata || !odata || index < 0 || index >= (int)ranked.size()) return LIBRETT_C_INVALID_PARAMETER;
    std::vector<BatchDesc> ds, fs;
    std::vector<long long> offs;
    if (!describe_problem(p, ranked[index], ds, fs, offs)) return LIBRETT_C_INTERNAL_ERROR;
    if (fallback) ds = fs;
    const long long ub = unit_bytes > 0 ? unit_bytes : kBatchUnitBytes;
    const size_t n = (size_t)p.vol * p.elem;
    const size_t g = 4096;
    std::vector<unsigned char> gin(n + 2 * g, 0xA5), gout(n + 2 * g, 0x5A);
    std::memcpy(gin.data() + g, idata, n);
    std::vector<int> start(ds.size() + 1, 0);
    for (size_t k = 0; k < ds.size(); k++) {
      batch_regroup(ds[k], ub);
      ds[k].in = gin.data() + g + offs[k];
      ds[k].out = gout.data() + g + offs[k];
      start[k + 1] = start[k] + ds[k].units;
    }
    const int gr = grid > 0 ? grid : dev.cu_count * 4;
    if (!replay_batch(ds.data(), start.data(), (int)ds.size(), start.back(), gr)) return LIBRETT_C_UNDEFINED_ERROR;
    for (size_t j = 0; j < g; j++)
      if (gout[j] != 0x5A || gout[g + n + j] != 0x5A) return LIBRETT_C_UNDEFINED_ERROR;
    std::memcpy(odata, gout.data() + g, n);
    return LIBRETT_C_SUCCESS;
  } catch (...) {
    return LIBRETT_C_INTERNAL_ERROR;
  }
}

}  // extern "C"
