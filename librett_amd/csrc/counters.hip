// Exact access counters of a candidate plan.
//
// The reference counts the global-memory transactions and cache lines of a
// plan with kernels that replay each transpose kernel's lane addresses and
// compare them across a warp (librettGpuModelKernel,
// src/GpuModelKernel.cpp:1013-1224; countTiled, countTiledCopy,
// countPacked, countPackedSplit at :343-946). The planner itself only
// estimates these counts from a few sampled work items
// (src/librett.cpp:183, src/plan.cpp:1073-1643).
//
// Here one workgroup takes one work item at a time (a tile for Tiled and
// TiledCopy, one staged round (chunk, outer index) for Packed and
// PackedSplit) and walks the item's wave instructions: each wave computes
// its 64 lane addresses the way the transpose kernel does, and counts the
// distinct 128-byte segments among the active lanes with ballots. One
// thread walks the item's contiguous runs in address order for the distinct
// lines and the partially written lines. Each workgroup writes its sums to
// its own slot (no atomics); the host adds the slots. host_counts runs the
// same per-item code on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "counters.h"
#include "devmem.h"
#include "kernel_ops.h"
#include "kernels.h"

namespace lrt {
namespace {

constexpr int kCountThreads = 256;
constexpr int kCountMaxGrid = 2048;

struct RunRef {
  int off, n;  // start and length in the run table (pairs of ints)
};

struct CountJob {
  int method;          // kTiled, kTiledCopy, kPacked or kPackedSplit
  int B;               // element bytes (Tiled, Packed) or word bytes (TiledCopy)
  int TX, TY, VX, VY;  // Tiled: tile and vector widths; TiledCopy: TX words, TY rows
  int nt, nreg;        // Packed: threads and registers (passes) per thread
  long long items;
  const int* runs;     // Packed: (start, length) element runs of each staged shape
  RunRef run[2][2];    // [variant][0 input side, 1 output side]
  union {
    TiledArgs tiled;
    CopyArgs copy;
    PackedArgs packed;
  } a;
};

__host__ __device__ inline int imin(int a, int b) { return a < b ? a : b; }

struct ItemCtx {
  int x0, y0, tx, ty;  // Tiled / TiledCopy: tile origin and extent
  int plus;            // Packed: staged shape variant
  long long in0, out0; // element (or word) offset of the item's base
};

// Tile origin of work item w; same order as TiledOps::item / CopyOps::item.
__host__ __device__ inline ItemCtx item_ctx(const CountJob& j, uint32_t w) {
  ItemCtx c{};
  if (j.method == kTiled) {
    const TiledArgs& a = j.a.tiled;
    const uint32_t mb = fdiv(w, a.ntiles_div);
    const uint32_t tile = w - mb * (uint32_t)a.ntiles;
    uint32_t tx, ty;
    if (a.group <= 1) {
      ty = fdiv(tile, a.ntx_div);
      tx = tile - ty * (uint32_t)a.ntx;
    } else {
      const uint32_t g = fdiv(tile, a.group_row_div);
      const uint32_t within = tile - g * a.group_row_div.d;
      const FastDiv& rows = (tile >= (uint32_t)a.last_group_start) ? a.last_div : a.group_div;
      tx = fdiv(within, rows);
      ty = g * (uint32_t)a.group + (within - tx * rows.d);
    }
    c.x0 = (int)tx * j.TX;
    c.y0 = (int)ty * j.TY;
    c.tx = imin(j.TX, a.X - c.x0);
    c.ty = imin(j.TY, a.Y - c.y0);
    int oi, oo;
    outer_offsets(a.outer, mb, oi, oo);
    c.in0 = oi;
    c.out0 = oo;
  } else if (j.method == kTiledCopy) {
    const CopyArgs& a = j.a.copy;
    const uint32_t mb = fdiv(w, a.ntiles_div);
    const uint32_t tile = w - mb * (uint32_t)a.ntiles;
    const uint32_t ty = fdiv(tile, a.ntx_div);
    const uint32_t tx = tile - ty * (uint32_t)a.ntx;
    c.x0 = (int)tx * j.TX;
    c.y0 = (int)ty * j.TY;
    c.tx = imin(j.TX, a.X - c.x0);
    c.ty = imin(j.TY, a.Y - c.y0);
    int oi, oo;
    outer_offsets(a.outer, mb, oi, oo);
    c.in0 = oi;
    c.out0 = oo;
  } else {
    // Round (split, o): chunk `split` of the split dim at outer index o, as
    // PackedOps::prepare and the outer loop of k_packed address it.
    const PackedArgs& a = j.a.packed;
    const int split = (int)(w / (uint32_t)a.vol_outer);
    const int o = (int)(w - (uint32_t)split * (uint32_t)a.vol_outer);
    const int q0 = a.split_dim / a.num_split;
    const int c0 = (int)((long long)split * a.split_dim / a.num_split);
    const int c1 = (int)((long long)(split + 1) * a.split_dim / a.num_split);
    c.plus = (c1 - c0) - q0;
    int oi, oo;
    outer_offsets(a.outer, (uint32_t)o, oi, oo);
    c.in0 = (long long)c0 * a.split_in_stride + oi;
    c.out0 = (long long)c0 * a.split_out_stride + oo;
  }
  return c;
}

__host__ __device__ inline int job_threads(const CountJob& j) {
  return (j.method == kPacked || j.method == kPackedSplit) ? j.nt : kCountThreads;
}

// Instructions (passes) per thread of one kind.
__host__ __device__ inline int job_passes(const CountJob& j, int store) {
  switch (j.method) {
    case kTiled:
      return store ? (j.TX * j.TY) / (j.VY * kCountThreads) : (j.TX * j.TY) / (j.VX * kCountThreads);
    case kTiledCopy:
      return j.TY / (kCountThreads / j.TX);
    default:
      return j.nreg;
  }
}

__host__ __device__ inline int job_access_bytes(const CountJob& j, int store) {
  if (j.method == kTiled) return j.B * (store ? j.VY : j.VX);
  return j.B;
}

// Byte offset of thread t's access in pass p of an item, false when the lane
// is inactive. Mirrors the lane arithmetic of TiledOps::load / store,
// CopyOps::run and PackedOps::prepare / load / store.
__host__ __device__ inline bool lane_access(const CountJob& j, const ItemCtx& c, int store, int p, int t,
                                            long long& off) {
  if (j.method == kTiled) {
    const TiledArgs& a = j.a.tiled;
    if (!store) {
      const int LR = j.TX / j.VX;
      const int RP = kCountThreads / LR;
      const int x = c.x0 + (t % LR) * j.VX;
      const int y = c.y0 + t / LR + p * RP;
      if (x >= a.X || y >= a.Y) return false;
      off = (long long)(c.in0 + x + (long long)tiled_in_y(a, (uint32_t)y)) * j.B;
    } else {
      const int LW = j.TY / j.VY;
      const int WP = kCountThreads / LW;
      const int y = c.y0 + (t % LW) * j.VY;
      const int x = c.x0 + t / LW + p * WP;
      if (x >= a.X || y >= a.Y) return false;
      off = (long long)(c.out0 + y + (long long)tiled_out_x(a, (uint32_t)x)) * j.B;
    }
    return true;
  }
  if (j.method == kTiledCopy) {
    const CopyArgs& a = j.a.copy;
    const int RP = kCountThreads / j.TX;
    const int x = c.x0 + t % j.TX;
    const int y = c.y0 + t / j.TX + p * RP;
    if (x >= a.X || y >= a.Y) return false;
    off = (long long)((store ? c.out0 : c.in0) + x + (long long)y * (store ? a.out_ys : a.in_ys)) * j.B;
    return true;
  }
  const StagedShape& s = j.a.packed.shape[c.plus];
  uint32_t k = (uint32_t)(t + p * j.nt);
  if ((int)k >= s.vol) return false;
  long long e = 0;
  for (int d = 0; d < s.n; d++) {
    const FastDiv& dv = store ? s.dout[d] : s.din[d];
    const uint32_t q = fdiv(k, dv);
    e += (long long)(k - q * dv.d) * (store ? s.ct_out[d] : s.ct_in[d]);
    k = q;
  }
  off = ((store ? c.out0 : c.in0) + e) * j.B;
  return true;
}

// Distinct 128-byte lines of runs given in ascending start order, merging
// runs that touch or overlap; `partial` counts lines that hold a boundary of
// a merged run that is not on a line boundary, i.e. lines the runs cover
// only in part.
struct LineWalk {
  long long s = -1, e = -1;
  long long last_line = -1, last_partial = -1;
  unsigned long long lines = 0, partial = 0;

  __host__ __device__ void add(long long rs, long long n) {
    if (n <= 0) return;
    if (e >= 0 && rs <= e) {
      if (rs + n > e) e = rs + n;
      return;
    }
    flush();
    s = rs;
    e = rs + n;
  }
  __host__ __device__ void flush() {
    if (e < 0) return;
    const long long first = s >> 7, last = (e - 1) >> 7;
    lines += (unsigned long long)(last - first + 1) - (first == last_line ? 1 : 0);
    if ((s & 127) && first != last_partial) {
      partial++;
      last_partial = first;
    }
    if ((e & 127) && last != last_partial) {
      partial++;
      last_partial = last;
    }
    last_line = last;
    s = e = -1;
  }
};

// Insertion sort of a few values.
__host__ __device__ inline void sort_small(long long* v, int n) {
  for (int i = 1; i < n; i++) {
    const long long x = v[i];
    int k = i - 1;
    while (k >= 0 && v[k] > x) {
      v[k + 1] = v[k];
      k--;
    }
    v[k + 1] = x;
  }
}

// Lines one item reads and writes: Tiled reads tile rows (y ascending) and
// writes tile columns (x ascending); TiledCopy reads and writes rows; Packed
// walks the staged shape's runs, which the host sorted by start.
__host__ __device__ inline void item_lines(const CountJob& j, const ItemCtx& c, unsigned long long* acc) {
  LineWalk rd, wr;
  const long long B = j.B;
  if (j.method == kTiled) {
    const TiledArgs& a = j.a.tiled;
    // Rows (columns) of a one-dim group are in address order; those of a
    // group of several dims are sorted first (a tile has at most 128).
    long long st[128];
    for (int y = 0; y < c.ty; y++) st[y] = (c.in0 + c.x0 + (long long)tiled_in_y(a, (uint32_t)(c.y0 + y))) * B;
    if (a.yg.n > 1) sort_small(st, c.ty);
    for (int y = 0; y < c.ty; y++) rd.add(st[y], c.tx * B);
    for (int x = 0; x < c.tx; x++) st[x] = (c.out0 + c.y0 + (long long)tiled_out_x(a, (uint32_t)(c.x0 + x))) * B;
    if (a.xg.n > 1) sort_small(st, c.tx);
    for (int x = 0; x < c.tx; x++) wr.add(st[x], c.ty * B);
  } else if (j.method == kTiledCopy) {
    const CopyArgs& a = j.a.copy;
    for (int y = 0; y < c.ty; y++) {
      rd.add((c.in0 + c.x0 + (long long)(c.y0 + y) * a.in_ys) * B, c.tx * B);
      wr.add((c.out0 + c.x0 + (long long)(c.y0 + y) * a.out_ys) * B, c.tx * B);
    }
  } else {
    const RunRef ri = j.run[c.plus][0], ro = j.run[c.plus][1];
    for (int r = 0; r < ri.n; r++) {
      const int* q = j.runs + 2 * (ri.off + r);
      rd.add((c.in0 + q[0]) * B, (long long)q[1] * B);
    }
    for (int r = 0; r < ro.n; r++) {
      const int* q = j.runs + 2 * (ro.off + r);
      wr.add((c.out0 + q[0]) * B, (long long)q[1] * B);
    }
  }
  rd.flush();
  wr.flush();
  acc[kReadLines] += rd.lines;
  acc[kWriteLines] += wr.lines;
  acc[kPartialLines] += wr.partial;
  acc[kItems] += 1;
}

// One wave instruction: requests, distinct 128-byte segments and bytes of
// its active lanes. An access of `bytes` may span two segments; both count.
__device__ inline void wave_count(bool act, long long off, int bytes, int lane, unsigned long long& req,
                                  unsigned long long& tran, unsigned long long& nbytes) {
  const unsigned long long m = __ballot(act);
  if (m == 0) return;
  const uint32_t lo = act ? (uint32_t)(off >> 7) : 0xffffffffu;
  const uint32_t hi = act ? (uint32_t)((off + bytes - 1) >> 7) : 0xffffffffu;
  int n = 0;
  unsigned long long left = m;
  while (left) {
    const int leader = __ffsll((unsigned long long)left) - 1;
    const uint32_t s = __shfl(lo, leader);
    left &= ~__ballot(act && lo == s);
    n++;
  }
  left = __ballot(act && hi != lo);
  while (left) {
    const int leader = __ffsll((unsigned long long)left) - 1;
    const uint32_t s = __shfl(hi, leader);
    if (__ballot(act && lo == s) == 0) n++;
    left &= ~__ballot(act && hi != lo && hi == s);
  }
  if (lane == 0) {
    req += 1;
    tran += (unsigned long long)n;
    nbytes += (unsigned long long)__popcll(m) * (unsigned long long)bytes;
  }
}

__global__ __launch_bounds__(kCountThreads) void k_count(CountJob job, unsigned long long* __restrict__ part) {
  __shared__ unsigned long long red[kCountThreads];
  unsigned long long acc[kNumCounts];
  for (int i = 0; i < kNumCounts; i++) acc[i] = 0;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int nthreads = job_threads(job);
  const int waves = (nthreads + 63) / 64;
  for (long long w = blockIdx.x; w < job.items; w += gridDim.x) {
    const ItemCtx c = item_ctx(job, (uint32_t)w);
    for (int store = 0; store < 2; store++) {
      const int np = job_passes(job, store);
      const int bytes = job_access_bytes(job, store);
      for (int p = 0; p < np; p++)
        for (int v = wv; v < waves; v += kCountThreads / 64) {
          const int t = v * 64 + lane;
          long long off = 0;
          const bool act = t < nthreads && lane_access(job, c, store, p, t, off);
          wave_count(act, off, bytes, lane, acc[kLdReq + store], acc[kLdTran + store], acc[kLdBytes + store]);
        }
    }
    if (threadIdx.x == 0) item_lines(job, c, acc);
  }
  for (int i = 0; i < kNumCounts; i++) {
    red[threadIdx.x] = acc[i];
    __syncthreads();
    for (int s = kCountThreads / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(size_t)blockIdx.x * kNumCounts + i] = red[0];
    __syncthreads();
  }
}

// Host version of wave_count over one wave's lanes.
void host_wave_count(const bool* act, const long long* off, int bytes, unsigned long long& req,
                     unsigned long long& tran, unsigned long long& nbytes) {
  long long segs[128];
  int ns = 0, na = 0;
  for (int l = 0; l < 64; l++) {
    if (!act[l]) continue;
    na++;
    segs[ns++] = off[l] >> 7;
    segs[ns++] = (off[l] + bytes - 1) >> 7;
  }
  if (na == 0) return;
  std::sort(segs, segs + ns);
  req += 1;
  tran += (unsigned long long)(std::unique(segs, segs + ns) - segs);
  nbytes += (unsigned long long)na * (unsigned long long)bytes;
}

// Fills a job from a candidate; `runs` receives the Packed run table (host
// memory), sorted by start within each (variant, side).
bool make_job(const Candidate& c, CountJob& j, std::vector<int>& runs) {
  std::memset(&j, 0, sizeof(j));
  j.method = c.method;
  runs.clear();
  switch (c.method) {
    case kTiled:
      j.B = c.elem;
      j.TX = c.tile_x;
      j.TY = c.tile_y;
      j.VX = c.vec_x;
      j.VY = c.vec_y;
      j.a.tiled = c.tiled;
      j.items = c.tiled.items;
      if (j.TX % j.VX || j.TY % j.VY || kCountThreads % (j.TX / j.VX) || kCountThreads % (j.TY / j.VY)) return false;
      return job_passes(j, 0) >= 1 && job_passes(j, 1) >= 1;
    case kTiledCopy:
      j.B = c.word;
      j.TX = c.copy_tx;
      if (j.TX <= 0 || kCountThreads % j.TX) return false;
      j.TY = (kCountThreads / j.TX) * 4;  // CopyOps<W, TX, 4>
      j.a.copy = c.copy;
      j.items = c.copy.items;
      return true;
    case kPacked:
    case kPackedSplit: {
      j.B = c.elem;
      j.nt = c.block;
      j.nreg = c.nreg;
      j.a.packed = c.packed;
      j.items = (long long)c.packed.num_split * c.packed.vol_outer;
      const int nvar = (c.method == kPackedSplit) ? 2 : 1;
      std::vector<std::pair<long long, int>> r;
      for (int v = 0; v < nvar; v++)
        for (int side = 0; side < 2; side++) {
          staged_runs(c.packed.shape[v], side == 1, r);
          std::sort(r.begin(), r.end());
          j.run[v][side].off = (int)(runs.size() / 2);
          j.run[v][side].n = (int)r.size();
          for (const auto& q : r) {
            runs.push_back((int)q.first);
            runs.push_back(q.second);
          }
        }
      return j.nt > 0 && j.nt <= 1024 && j.nreg > 0;
    }
    default:
      return false;
  }
}

// Calls f(item, store, pass, thread, offset) for every active access of the
// enumeration, and returns the per-instruction counts through `acc`.
template <class F>
void host_walk(const CountJob& j, unsigned long long* acc, F&& f) {
  const int nthreads = job_threads(j);
  const int waves = (nthreads + 63) / 64;
  bool act[64];
  long long off[64];
  for (long long w = 0; w < j.items; w++) {
    const ItemCtx c = item_ctx(j, (uint32_t)w);
    for (int store = 0; store < 2; store++) {
      const int np = job_passes(j, store);
      const int bytes = job_access_bytes(j, store);
      for (int p = 0; p < np; p++)
        for (int v = 0; v < waves; v++) {
          for (int l = 0; l < 64; l++) {
            const int t = v * 64 + l;
            off[l] = 0;
            act[l] = t < nthreads && lane_access(j, c, store, p, t, off[l]);
            if (act[l]) f(w, store, p, t, off[l], bytes);
          }
          host_wave_count(act, off, bytes, acc[kLdReq + store], acc[kLdTran + store], acc[kLdBytes + store]);
        }
    }
    item_lines(j, c, acc);
  }
}

}  // namespace

hipError_t device_counts(const Candidate& c, hipStream_t stream, AccessCounts& out) {
  std::memset(&out, 0, sizeof(out));
  CountJob job;
  std::vector<int> runs;
  if (!make_job(c, job, runs)) return hipErrorInvalidValue;
  if (job.items <= 0) return hipSuccess;
  const int grid = (int)std::min<long long>(job.items, kCountMaxGrid);
  unsigned long long* d_part = nullptr;
  int* d_runs = nullptr;
  std::vector<unsigned long long> h((size_t)grid * kNumCounts);
  hipError_t e = dev_malloc((void**)&d_part, h.size() * sizeof(unsigned long long));
  if (e == hipSuccess && !runs.empty()) {
    e = dev_malloc((void**)&d_runs, runs.size() * sizeof(int));
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_runs, runs.data(), runs.size() * sizeof(int), hipMemcpyHostToDevice, stream);
  }
  if (e == hipSuccess) {
    job.runs = d_runs;
    void* args[2] = {&job, &d_part};
    e = hipLaunchKernel(reinterpret_cast<const void*>(&k_count), dim3(grid), dim3(kCountThreads), args, 0, stream);
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), d_part, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (d_part) (void)dev_free(d_part);
  if (d_runs) (void)dev_free(d_runs);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return e;
  }
  for (int b = 0; b < grid; b++)
    for (int i = 0; i < kNumCounts; i++) out.v[i] += h[(size_t)b * kNumCounts + i];
  return hipSuccess;
}

bool host_counts(const Candidate& c, AccessCounts& out) {
  std::memset(&out, 0, sizeof(out));
  CountJob job;
  std::vector<int> runs;
  if (!make_job(c, job, runs)) return false;
  job.runs = runs.data();
  host_walk(job, out.v, [](long long, int, int, int, long long, int) {});
  return true;
}

bool trace_enumeration(const Candidate& c, std::vector<AccessRecord>& recs) {
  recs.clear();
  CountJob job;
  std::vector<int> runs;
  if (!make_job(c, job, runs)) return false;
  job.runs = runs.data();
  unsigned long long acc[kNumCounts] = {};
  host_walk(job, acc, [&](long long w, int store, int p, int t, long long off, int bytes) {
    recs.push_back({w, store, p, t, off, bytes});
  });
  return true;
}

namespace {

struct TraceObserver : ReplayObserver {
  const unsigned char* in;
  const unsigned char* out;
  size_t n;
  long long item = -1;
  int thread = -1;
  int seq[2] = {0, 0};
  bool stray = false;
  std::vector<AccessRecord>* recs;
  void at(long long i, int t) override {
    if (i != item || t != thread) seq[0] = seq[1] = 0;
    item = i;
    thread = t;
  }
  void access(const void* p, int bytes, int store) override {
    const unsigned char* base = store ? out : in;
    const unsigned char* q = static_cast<const unsigned char*>(p);
    if (q < base || q + bytes > base + n) stray = true;
    recs->push_back({item, store, seq[store]++, thread, (long long)(q - base), bytes});
  }
};

}  // namespace

bool trace_kernel(const Candidate& c, std::vector<AccessRecord>& recs) {
  recs.clear();
  if (c.method == kTrivial) return false;
  const size_t n = (size_t)(c.bytes / 2);
  std::vector<unsigned char> in(n + 16, 0), out(n + 16, 0);
  // 16-byte aligned bases, as device allocations are (vector accesses).
  unsigned char* pi = in.data() + ((16 - ((uintptr_t)in.data() & 15)) & 15);
  unsigned char* po = out.data() + ((16 - ((uintptr_t)out.data() & 15)) & 15);
  if ((size_t)(pi - in.data()) + n > in.size() || (size_t)(po - out.data()) + n > out.size()) return false;
  TraceObserver obs;
  obs.in = pi;
  obs.out = po;
  obs.n = n;
  obs.recs = &recs;
  g_replay_observer = &obs;
  const bool ok = replay_candidate(c, pi, po);
  g_replay_observer = nullptr;
  return ok && !obs.stray;
}

}  // namespace lrt
