// Serving engine: continuous batching of transpose requests on one GPU.
//
// Requests (a problem plus device input/output pointers) are queued by any
// thread. One dispatcher thread owns a few launch slots, each a stream with
// its own pinned staging and device descriptor area. Whenever a slot is
// free, every queued request (up to a cap) becomes one micro-batch: its
// descriptors are copied to the device and the batched engine's persistent
// kernel runs them; micro-batches on different slots overlap on the device,
// so requests that arrive while others run are admitted as soon as a slot
// frees up. Completion is observed with one event per micro-batch.
//
// Output buffers come from a device arena with a hard byte budget:
// admission waits while the arena cannot hold a request (a stall), and every
// byte admitted is returned when its request completes. Plans are kept in an
// LRU cache keyed by (dims, permutation, element size).
//
// A native Poisson load generator drives the engine for measurements.
// The reference has no serving mode; its per-request path is librettPlan +
// librettExecute (src/librett.cpp:110-222, 372-385).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "arena.h"
#include "devmem.h"
#include "../../include/librett_c.h"
#include "batch.h"
#include "lru_cache.h"
#include "planner.h"

namespace lrt {
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit arena over one device allocation, 256-byte granules, with
// coalescing of free neighbours.
// The serving engine's device arena: one allocation from the device
// allocator, handed out in ranges by a RangeAllocator.
class Arena {
 public:
  bool init(size_t bytes) {
    const size_t cap = (bytes + RangeAllocator::kGranule - 1) / RangeAllocator::kGranule * RangeAllocator::kGranule;
    if (dev_malloc(&base_, cap) != hipSuccess) {
      (void)hipGetLastError();
      base_ = nullptr;
      return false;
    }
    ranges_.reset(cap);
    return true;
  }
  ~Arena() {
    if (base_) (void)dev_free(base_);
  }
  void* alloc(size_t bytes) {
    const size_t off = ranges_.alloc(bytes);
    return off == RangeAllocator::kNone ? nullptr : (char*)base_ + off;
  }
  bool release(void* p) {
    if (!p || !owns(p)) return false;
    return ranges_.release((size_t)((char*)p - (char*)base_));
  }
  size_t in_use() { return ranges_.in_use(); }
  size_t peak() { return ranges_.peak(); }
  // Restarts the peak at the bytes in use now.
  void reset_peak() { ranges_.reset_peak(); }
  size_t capacity() const { return ranges_.capacity(); }
  bool owns(const void* p) const { return p >= base_ && p < (const char*)base_ + ranges_.capacity(); }

 private:
  void* base_ = nullptr;
  RangeAllocator ranges_;
};

struct Request {
  unsigned long long id = 0;
  Problem prob;
  Candidate plan;
  const void* in = nullptr;
  void* out = nullptr;
  bool release_out = false;  // generator requests return their output to the arena
  bool release_in = false;   // ... and their input, when it came from the arena too
  double t_arrival = 0, t_done = 0;
  int status = 0;            // 0 queued, 1 running, 2 done, <0 failed
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<std::shared_ptr<Request>> reqs;
  BatchDesc* h_desc = nullptr;  // pinned staging
  int* h_start = nullptr;
  BatchDesc* d_desc = nullptr;
  int* d_start = nullptr;
  int cap = 0;                  // descriptors
};

class Server {
 public:
  ~Server() { shutdown(); }

  int init(size_t arena_bytes, int nslots, int max_batch) {
    if (hipGetDevice(&device_) != hipSuccess) return LIBRETT_C_INTERNAL_ERROR;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device_) != hipSuccess) return LIBRETT_C_INTERNAL_ERROR;
    dev_.device_id = device_;
    dev_.cu_count = prop.multiProcessorCount;
    dev_.lds_per_block = (int)prop.sharedMemPerBlock;
    dev_.lds_per_cu = (int)std::max<size_t>(prop.sharedMemPerBlock, prop.maxSharedMemoryPerMultiProcessor);
    dev_.max_threads_per_block = prop.maxThreadsPerBlock;
    dev_.warp_size = prop.warpSize;
    dev_.max_waves_per_cu = std::max(1, prop.maxThreadsPerMultiProcessor / std::max(1, prop.warpSize));
    if (prop.l2CacheSize > 0) dev_.l2_bytes = (double)prop.l2CacheSize * std::max(1, dev_.cu_count / 32);
    dev_.has_gpu = 1;
    grid_max_ = 0;
    for (int f = 0; f < kBatchFamilies; f++) {
      grid_[f] = dev_.cu_count * std::max(1, batch_occupancy(f));
      grid_max_ = std::max(grid_max_, grid_[f]);
    }
    max_batch_ = std::max(1, max_batch);
    if (!arena_.init(arena_bytes)) return LIBRETT_C_INTERNAL_ERROR;
    slots_.resize(std::max(1, nslots));
    for (auto& s : slots_) {
      s.cap = 2 * max_batch_ + 2;
      if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
          hipEventCreateWithFlags(&s.done, hipEventDisableTiming) != hipSuccess ||
          hipHostMalloc((void**)&s.h_desc, s.cap * sizeof(BatchDesc), hipHostMallocDefault) != hipSuccess ||
          hipHostMalloc((void**)&s.h_start, (s.cap + kBatchFamilies) * sizeof(int), hipHostMallocDefault) !=
              hipSuccess ||
          dev_malloc((void**)&s.d_desc, s.cap * sizeof(BatchDesc)) != hipSuccess ||
          dev_malloc((void**)&s.d_start, (s.cap + kBatchFamilies) * sizeof(int)) != hipSuccess) {
        (void)hipGetLastError();
        return LIBRETT_C_INTERNAL_ERROR;
      }
    }
    stop_ = false;
    dispatcher_ = std::thread([this] { dispatch_loop(); });
    return LIBRETT_C_SUCCESS;
  }

  void shutdown() {
    if (dispatcher_.joinable()) {
      {
        std::lock_guard<std::mutex> lock(mu_);
        stop_ = true;
      }
      cv_.notify_all();
      dispatcher_.join();
    }
    for (auto& s : slots_) {
      if (s.stream) (void)hipStreamSynchronize(s.stream);
      if (s.done) (void)hipEventDestroy(s.done);
      if (s.stream) (void)hipStreamDestroy(s.stream);
      if (s.h_desc) (void)hipHostFree(s.h_desc);
      if (s.h_start) (void)hipHostFree(s.h_start);
      if (s.d_desc) (void)dev_free(s.d_desc);
      if (s.d_start) (void)dev_free(s.d_start);
      s = Slot();
    }
  }

  Arena& arena() { return arena_; }

  // Plan of a problem from the LRU cache, planned on a miss (in the calling
  // thread; the time misses take is part of the statistics).
  bool plan_for(const Problem& p, Candidate& c) {
    const std::string key = key_of(p);
    if (plans_.get(key, c)) return true;
    const double t0 = now_s();
    std::vector<Candidate> cands = batch_candidates(p, dev_);
    if (cands.empty()) return false;
    c = cands[0];
    plans_.put(key, c);
    const double dt = now_s() - t0;
    std::lock_guard<std::mutex> lock(mu_);
    plan_misses_++;
    plan_s_ += dt;
    return true;
  }

  int submit(const Problem& p, const void* in, void* out, bool release_out, double t_arrival,
             unsigned long long* ticket, bool release_in = false) {
    auto r = std::make_shared<Request>();
    r->prob = p;
    if (!plan_for(p, r->plan)) return LIBRETT_C_INTERNAL_ERROR;
    r->in = in;
    r->out = out;
    r->release_out = release_out;
    r->release_in = release_in;
    r->t_arrival = t_arrival > 0 ? t_arrival : now_s();
    {
      std::lock_guard<std::mutex> lock(mu_);
      r->id = next_id_++;
      pending_.push_back(r);
      live_[r->id] = r;
      submitted_++;
    }
    cv_.notify_all();
    if (ticket) *ticket = r->id;
    return LIBRETT_C_SUCCESS;
  }

  int wait(unsigned long long id, double* latency) {
    std::unique_lock<std::mutex> lock(mu_);
    auto it = live_.find(id);
    if (it == live_.end()) {
      auto d = finished_.find(id);
      if (d == finished_.end()) return LIBRETT_C_INVALID_PARAMETER;
      if (latency) *latency = d->second;
      return LIBRETT_C_SUCCESS;
    }
    std::shared_ptr<Request> r = it->second;
    done_cv_.wait(lock, [&] { return r->status == 2 || r->status < 0; });
    if (latency) *latency = r->t_done - r->t_arrival;
    return r->status == 2 ? LIBRETT_C_SUCCESS : LIBRETT_C_INTERNAL_ERROR;
  }

  void drain() {
    std::unique_lock<std::mutex> lock(mu_);
    done_cv_.wait(lock, [&] { return live_.empty(); });
  }

  void stats(librett_c_server_stats* st) {
    std::lock_guard<std::mutex> lock(mu_);
    std::memset(st, 0, sizeof(*st));
    st->submitted = submitted_;
    st->completed = (long long)lat_.size();
    st->failed = failed_;
    st->bytes = bytes_;
    st->batches = batches_;
    st->mean_batch = batches_ ? (double)batched_reqs_ / batches_ : 0;
    st->admission_stalls = stalls_;
    st->stall_seconds = stall_s_;
    st->arena_peak_bytes = (long long)arena_.peak();
    st->arena_in_use_bytes = (long long)arena_.in_use();
    if (!lat_.empty()) {
      std::vector<double> v = lat_;
      std::sort(v.begin(), v.end());
      auto q = [&](double f) { return v[std::min(v.size() - 1, (size_t)(f * (v.size() - 1) + 0.5))]; };
      double sum = 0;
      for (double x : v) sum += x;
      st->p50_us = q(0.50) * 1e6;
      st->p90_us = q(0.90) * 1e6;
      st->p99_us = q(0.99) * 1e6;
      st->max_us = v.back() * 1e6;
      st->mean_us = sum / v.size() * 1e6;
      st->wall_seconds = last_done_ - first_arrival_;
      st->gbs = st->wall_seconds > 0 ? (double)bytes_ / st->wall_seconds / 1e9 : 0;
    }
    st->plan_misses = plan_misses_;
    st->plan_seconds = plan_s_;
  }

  void reset_stats() {
    std::lock_guard<std::mutex> lock(mu_);
    lat_.clear();
    bytes_ = 0;
    batches_ = batched_reqs_ = 0;
    stalls_ = 0;
    stall_s_ = 0;
    submitted_ = failed_ = 0;
    plan_misses_ = 0;
    plan_s_ = 0;
    first_arrival_ = 1e300;
    last_done_ = 0;
    finished_.clear();
    finished_order_.clear();
    arena_.reset_peak();
  }

  void add_stall(double s) {
    std::lock_guard<std::mutex> lock(mu_);
    stalls_++;
    stall_s_ += s;
  }

  // Waits (bounded) until a request completes; used by the generator while
  // the arena is full.
  void wait_any_completion(double timeout_s) {
    std::unique_lock<std::mutex> lock(mu_);
    const unsigned long long seen = completions_;
    done_cv_.wait_for(lock, std::chrono::duration<double>(timeout_s), [&] { return completions_ != seen; });
  }

 private:
  static std::string key_of(const Problem& p) {
    std::string k = std::to_string(p.elem) + ":";
    for (int i = 0; i < p.rank; i++) k += std::to_string(p.dim[i]) + ",";
    k += "/";
    for (int i = 0; i < p.rank; i++) k += std::to_string(p.perm[i]) + ",";
    return k;
  }

  // Fills a slot with queued requests and launches them. Called with mu_
  // held (through `lock`); the requests are taken off the queue under the
  // lock, but describing them, the uploads and the launches run without it,
  // so clients can keep submitting meanwhile. Only the dispatcher touches a
  // slot. Requests that do not fit the slot go back to the front of the
  // queue in their order.
  bool fill_and_launch(Slot& s, std::unique_lock<std::mutex>& lock) {
    std::vector<std::shared_ptr<Request>> cand;
    while (!pending_.empty() && (int)cand.size() < max_batch_) {
      cand.push_back(pending_.front());
      pending_.pop_front();
    }
    if (cand.empty()) return false;
    for (auto& r : cand) r->status = 1;
    lock.unlock();

    std::vector<std::shared_ptr<Request>> take, failed, back;
    int ndesc = 0;
    std::vector<BatchDesc> ds, fs;
    std::vector<long long> offs;
    std::vector<BatchDesc> use;
    for (size_t i = 0; i < cand.size(); i++) {
      const std::shared_ptr<Request>& r = cand[i];
      if (!describe_problem(r->prob, r->plan, ds, fs, offs)) {
        failed.push_back(r);
        continue;
      }
      if (ndesc + (int)ds.size() > s.cap) {
        back.assign(cand.begin() + (long)i, cand.end());
        break;
      }
      const int align = batch_alignment(r->prob, r->plan);
      for (size_t k = 0; k < ds.size(); k++) {
        const char* in = (const char*)r->in + offs[k];
        char* out = (char*)r->out + offs[k];
        const bool mis = (((uintptr_t)in | (uintptr_t)out) & (uintptr_t)(align - 1)) != 0;
        BatchDesc d = mis ? fs[k] : ds[k];
        d.in = in;
        d.out = out;
        use.push_back(d);
        ndesc++;
      }
      take.push_back(r);
    }
    bool ok = true;
    if (!take.empty()) {
      long long bytes = 0;
      for (auto& r : take) bytes += 2LL * r->prob.vol * r->prob.elem;
      const long long unit_bytes = batch_unit_bytes(bytes, grid_max_);
      for (int k = 0; k < ndesc; k++) batch_regroup(use[k], unit_bytes);
      // One launch per kernel family, in family order on the slot's stream.
      BatchLayout L;
      batch_layout(use, L, nullptr);
      bool fits = true;
      for (int f = 0; f < kBatchFamilies; f++) fits = fits && L.total[f] < (1LL << 31);
      for (int k = 0; k < ndesc; k++) s.h_desc[k] = use[k];
      for (size_t k = 0; k < L.start.size(); k++) s.h_start[k] = L.start[k];
      ok = fits &&
           hipMemcpyAsync(s.d_desc, s.h_desc, ndesc * sizeof(BatchDesc), hipMemcpyHostToDevice, s.stream) ==
               hipSuccess &&
           hipMemcpyAsync(s.d_start, s.h_start, L.start.size() * sizeof(int), hipMemcpyHostToDevice, s.stream) ==
               hipSuccess &&
           launch_batch_all(s.d_desc, s.d_start, L, grid_, s.stream) == hipSuccess &&
           hipEventRecord(s.done, s.stream) == hipSuccess;
      if (!ok) (void)hipGetLastError();
    }

    lock.lock();
    for (auto it = back.rbegin(); it != back.rend(); ++it) {
      (*it)->status = 0;
      pending_.push_front(*it);
    }
    if (!ok) failed.insert(failed.end(), take.begin(), take.end());
    for (auto& r : failed) {
      r->status = -1;
      failed_++;
      live_.erase(r->id);
    }
    if (!failed.empty()) done_cv_.notify_all();
    if (!ok || take.empty()) return !failed.empty();
    s.reqs = std::move(take);
    s.busy = true;
    batches_++;
    batched_reqs_ += s.reqs.size();
    return true;
  }

  void complete(Slot& s, double t) {
    for (auto& r : s.reqs) {
      r->t_done = t;
      r->status = 2;
      lat_.push_back(t - r->t_arrival);
      bytes_ += 2LL * r->prob.vol * r->prob.elem;
      first_arrival_ = std::min(first_arrival_, r->t_arrival);
      last_done_ = std::max(last_done_, t);
      if (r->release_out) arena_.release(r->out);
      if (r->release_in) arena_.release(const_cast<void*>(r->in));
      // Latencies of finished requests stay available to wait() for the
      // most recent kFinishedKeep of them.
      finished_[r->id] = t - r->t_arrival;
      finished_order_.push_back(r->id);
      if (finished_order_.size() > kFinishedKeep) {
        finished_.erase(finished_order_.front());
        finished_order_.pop_front();
      }
      live_.erase(r->id);
    }
    completions_ += s.reqs.size();
    s.reqs.clear();
    s.busy = false;
  }

  void dispatch_loop() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lock(mu_);
    while (true) {
      bool any_busy = false, progressed = false;
      for (auto& s : slots_) {
        if (s.busy) {
          lock.unlock();
          const hipError_t q = hipEventQuery(s.done);
          lock.lock();
          if (q == hipSuccess) {
            complete(s, now_s());
            progressed = true;
          } else if (q != hipErrorNotReady) {
            (void)hipGetLastError();
            for (auto& r : s.reqs) {
              r->status = -1;
              failed_++;
              live_.erase(r->id);
            }
            s.reqs.clear();
            s.busy = false;
            progressed = true;
          }
        }
        if (!s.busy && !pending_.empty()) progressed |= fill_and_launch(s, lock);
        any_busy |= s.busy;
      }
      if (progressed) done_cv_.notify_all();
      if (stop_ && !any_busy && pending_.empty()) break;
      if (!any_busy && pending_.empty()) {
        cv_.wait(lock, [&] { return stop_ || !pending_.empty(); });
      } else if (!progressed) {
        // Work in flight: poll again shortly without holding the lock.
        lock.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(5));
        lock.lock();
      }
    }
  }

  int device_ = 0;
  DeviceInfo dev_;
  int grid_[kBatchFamilies] = {};
  int grid_max_ = 0;
  int max_batch_ = 64;
  Arena arena_;
  std::vector<Slot> slots_;
  LruCache<std::string, Candidate> plans_{4096};

  std::mutex mu_;
  std::condition_variable cv_, done_cv_;
  std::deque<std::shared_ptr<Request>> pending_;
  std::unordered_map<unsigned long long, std::shared_ptr<Request>> live_;
  static constexpr size_t kFinishedKeep = 1 << 20;
  std::unordered_map<unsigned long long, double> finished_;
  std::deque<unsigned long long> finished_order_;
  std::thread dispatcher_;
  bool stop_ = true;
  unsigned long long next_id_ = 1, completions_ = 0;
  long long submitted_ = 0, failed_ = 0, bytes_ = 0, batches_ = 0, batched_reqs_ = 0, stalls_ = 0;
  long long plan_misses_ = 0;
  double stall_s_ = 0, plan_s_ = 0, first_arrival_ = 1e300, last_done_ = 0;
  std::vector<double> lat_;
};

std::mutex g_srv_mu;
std::unordered_map<unsigned int, std::shared_ptr<Server>> g_servers;
std::atomic<unsigned int> g_next_srv{1};

std::shared_ptr<Server> find_server(unsigned int h) {
  std::lock_guard<std::mutex> lock(g_srv_mu);
  auto it = g_servers.find(h);
  return it == g_servers.end() ? nullptr : it->second;
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

// A random problem of about `bytes` bytes: rank in [rlo, rhi], dims drawn
// log-uniformly and rounded so the volume lands near the target, a random
// permutation other than the identity.
void random_problem(std::mt19937_64& g, double bytes, int rlo, int rhi, int elem, std::vector<int>& dim,
                    std::vector<int>& perm) {
  std::uniform_int_distribution<int> rd(rlo, rhi);
  const int rank = rd(g);
  const double vol = std::max(2.0, bytes / elem);
  std::uniform_real_distribution<double> u(0.3, 1.0);
  std::vector<double> w(rank);
  double sw = 0;
  for (auto& x : w) sw += (x = u(g));
  dim.assign(rank, 1);
  long long prod = 1;
  for (int i = 0; i < rank; i++) {
    dim[i] = std::max(1, (int)std::lround(std::pow(vol, w[i] / sw)));
    prod *= dim[i];
  }
  int j = (int)(std::max_element(dim.begin(), dim.end()) - dim.begin());
  dim[j] = std::max(1, (int)std::lround(dim[j] * vol / (double)prod));
  perm.resize(rank);
  for (int i = 0; i < rank; i++) perm[i] = i;
  for (int tries = 0; tries < 8; tries++) {
    std::shuffle(perm.begin(), perm.end(), g);
    bool ident = true;
    for (int i = 0; i < rank; i++) ident &= perm[i] == i;
    if (!ident || rank == 1) break;
  }
}

}  // namespace
}  // namespace lrt

using namespace lrt;

extern "C" {

int librett_c_server_create(unsigned int* handle, size_t arena_bytes, int slots, int max_batch) {
  if (!handle || arena_bytes < 4096 || slots < 1 || max_batch < 1) return LIBRETT_C_INVALID_PARAMETER;
  try {
    auto s = std::make_shared<Server>();
    const int r = s->init(arena_bytes, slots, max_batch);
    if (r != LIBRETT_C_SUCCESS) return r;
    const unsigned int h = g_next_srv.fetch_add(1);
    std::lock_guard<std::mutex> lock(g_srv_mu);
    g_servers[h] = s;
    *handle = h;
    return LIBRETT_C_SUCCESS;
  } catch (...) {
    return LIBRETT_C_INTERNAL_ERROR;
  }
}

int librett_c_server_alloc(unsigned int handle, size_t bytes, void** ptr) {
  auto s = find_server(handle);
  if (!s) return LIBRETT_C_INVALID_PLAN;
  if (!ptr || bytes == 0) return LIBRETT_C_INVALID_PARAMETER;
  *ptr = s->arena().alloc(bytes);
  return *ptr ? LIBRETT_C_SUCCESS : LIBRETT_C_BUSY;
}

int librett_c_server_free(unsigned int handle, void* ptr) {
  auto s = find_server(handle);
  if (!s) return LIBRETT_C_INVALID_PLAN;
  return s->arena().release(ptr) ? LIBRETT_C_SUCCESS : LIBRETT_C_INVALID_PARAMETER;
}

int librett_c_server_submit(unsigned int handle, int rank, const int* dim, const int* perm, size_t elem,
                            const void* idata, void* odata, unsigned long long* ticket) {
  auto s = find_server(handle);
  if (!s) return LIBRETT_C_INVALID_PLAN;
  int r = check_problem(rank, dim, perm, elem);
  if (r != LIBRETT_C_SUCCESS) return r;
  if (!idata || !odata || idata == odata) return LIBRETT_C_INVALID_PARAMETER;
  try {
    return s->submit(canonicalize(rank, dim, perm, (int)elem), idata, odata, false, 0, ticket);
  } catch (...) {
    return LIBRETT_C_INTERNAL_ERROR;
  }
}

int librett_c_server_wait(unsigned int handle, unsigned long long ticket, double* latency_s) {
  auto s = find_server(handle);
  if (!s) return LIBRETT_C_INVALID_PLAN;
  return s->wait(ticket, latency_s);
}

int librett_c_server_drain(unsigned int handle) {
  auto s = find_server(handle);
  if (!s) return LIBRETT_C_INVALID_PLAN;
  s->drain();
  return LIBRETT_C_SUCCESS;
}

int librett_c_server_stats_get(unsigned int handle, librett_c_server_stats* st) {
  auto s = find_server(handle);
  if (!s) return LIBRETT_C_INVALID_PLAN;
  if (!st) return LIBRETT_C_INVALID_PARAMETER;
  s->stats(st);
  return LIBRETT_C_SUCCESS;
}

int librett_c_server_reset_stats(unsigned int handle) {
  auto s = find_server(handle);
  if (!s) return LIBRETT_C_INVALID_PLAN;
  s->reset_stats();
  return LIBRETT_C_SUCCESS;
}

int librett_c_server_destroy(unsigned int handle) {
  std::shared_ptr<Server> s;
  {
    std::lock_guard<std::mutex> lock(g_srv_mu);
    auto it = g_servers.find(handle);
    if (it == g_servers.end()) return LIBRETT_C_INVALID_PLAN;
    s = it->second;
    g_servers.erase(it);
  }
  s->drain();
  s->shutdown();
  return LIBRETT_C_SUCCESS;
}

int librett_c_server_run_poisson(unsigned int handle, const librett_c_poisson* cfg, const void* input_pool,
                                 size_t pool_bytes) {
  auto s = find_server(handle);
  if (!s) return LIBRETT_C_INVALID_PLAN;
  if (!cfg || !input_pool || cfg->qps <= 0 || cfg->requests < 1 || (!cfg->fresh && cfg->catalog < 1) ||
      cfg->min_bytes <= 0 || cfg->max_bytes < cfg->min_bytes || cfg->rank_lo < 1 || cfg->rank_hi < cfg->rank_lo ||
      cfg->nelems < 1 || cfg->nelems > 3 || (size_t)cfg->max_bytes > pool_bytes)
    return LIBRETT_C_INVALID_PARAMETER;
  for (int i = 0; i < cfg->nelems; i++)
    if (cfg->elems[i] != 4 && cfg->elems[i] != 8 && cfg->elems[i] != 16) return LIBRETT_C_INVALID_PARAMETER;
  try {
    // A request shape of log-uniform bytes in [min, max]; false if it does
    // not fit the pool or the arena (the caller draws again).
    std::uniform_real_distribution<double> lu(std::log(cfg->min_bytes), std::log(cfg->max_bytes));
    std::uniform_int_distribution<int> pe(0, cfg->nelems - 1);
    const size_t room = cfg->inputs_in_arena ? s->arena().capacity() / 2 : s->arena().capacity();
    auto draw = [&](std::mt19937_64& g, Problem& p) {
      const int elem = cfg->elems[pe(g)];
      std::vector<int> dim, perm;
      random_problem(g, std::exp(lu(g)), cfg->rank_lo, cfg->rank_hi, elem, dim, perm);
      if (check_problem((int)dim.size(), dim.data(), perm.data(), elem) != LIBRETT_C_SUCCESS) return false;
      p = canonicalize((int)dim.size(), dim.data(), perm.data(), elem);
      const size_t bytes = (size_t)p.vol * p.elem;
      return bytes <= pool_bytes && bytes <= room;
    };
    // Catalog of request shapes, planned up front so the plan cache is warm.
    std::vector<Problem> catalog;
    if (!cfg->fresh) {
      std::mt19937_64 g(cfg->seed);
      for (int i = 0; i < cfg->catalog; i++) {
        Problem p;
        Candidate c;
        if (draw(g, p) && s->plan_for(p, c)) catalog.push_back(p);
      }
      if (catalog.empty()) return LIBRETT_C_INTERNAL_ERROR;
    }
    const int clients = std::max(1, std::min(cfg->clients, 64));
    std::atomic<int> failed_rc{LIBRETT_C_SUCCESS};
    const double t0 = now_s();
    // Client k: its own Poisson stream of rate qps / clients and its share of
    // the requests; shapes from the catalog or drawn fresh (then planned in
    // submit, on this thread).
    auto client = [&](int k) {
      std::mt19937_64 g(cfg->seed * 7919 + 104729ULL * (k + 1));
      std::exponential_distribution<double> gap(cfg->qps / clients);
      const int n = cfg->requests / clients + (k < cfg->requests % clients ? 1 : 0);
      double t_next = t0;
      for (int i = 0; i < n && failed_rc.load() == LIBRETT_C_SUCCESS; i++) {
        t_next += gap(g);
        Problem p;
        if (cfg->fresh) {
          int tries = 0;
          while (!draw(g, p))
            if (++tries > 1000) {
              failed_rc = LIBRETT_C_INTERNAL_ERROR;
              return;
            }
        } else {
          p = catalog[std::uniform_int_distribution<size_t>(0, catalog.size() - 1)(g)];
        }
        const size_t bytes = (size_t)p.vol * p.elem;
        // Pool slice for the input unless it comes from the arena.
        const size_t nslices = (pool_bytes - bytes) / 256 + 1;
        const char* pool_in =
            (const char*)input_pool + (std::uniform_int_distribution<size_t>(0, nslices - 1)(g)) * 256;
        while (now_s() < t_next) std::this_thread::yield();
        // Admission: room for the output (and the input) in the arena, all
        // or nothing, so waiting clients hold no partial reservations.
        void* out = nullptr;
        void* in = nullptr;
        auto admit = [&] {
          out = s->arena().alloc(bytes);
          if (!out) return false;
          if (!cfg->inputs_in_arena) return true;
          in = s->arena().alloc(bytes);
          if (in) return true;
          s->arena().release(out);
          out = nullptr;
          return false;
        };
        if (!admit()) {
          const double ts = now_s();
          while (!admit()) s->wait_any_completion(1e-3);
          s->add_stall(now_s() - ts);
        }
        const int r = s->submit(p, cfg->inputs_in_arena ? in : pool_in, out, true, t_next, nullptr,
                                cfg->inputs_in_arena != 0);
        if (r != LIBRETT_C_SUCCESS) {
          s->arena().release(out);
          if (in) s->arena().release(in);
          failed_rc = r;
          return;
        }
      }
    };
    if (clients == 1) {
      client(0);
    } else {
      std::vector<std::thread> ts;
      for (int k = 0; k < clients; k++) ts.emplace_back(client, k);
      for (auto& t : ts) t.join();
    }
    s->drain();
    return failed_rc.load();
  } catch (...) {
    return LIBRETT_C_INTERNAL_ERROR;
  }
}

}  // extern "C"
