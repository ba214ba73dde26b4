// Host planner for the gfx950 transpose library. See planner.h.
#include "planner.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>

#include "kernels.h"

namespace lrt {

namespace {

// Blocks per CU that Tiled / TiledCopy grids are capped at (each block loops
// over the remaining tiles). LIBRETT_GRID_PER_CU overrides it.
int grid_per_cu() {
  static int v = [] {
    const char* e = std::getenv("LIBRETT_GRID_PER_CU");
    int x = e ? std::atoi(e) : 0;
    return x > 0 ? x : 64;
  }();
  return v;
}

// Tile-row group of the Tiled order (16 measured best for 256 MiB f32 and
// f64 2-D transposes on MI355X: +20-25% over row-major tile order in a
// steady stream of transposes); LIBRETT_TILE_GROUP overrides it.
int tile_group() {
  static int v = [] {
    const char* e = std::getenv("LIBRETT_TILE_GROUP");
    int x = e ? std::atoi(e) : 0;
    return x > 0 ? x : 16;
  }();
  return v;
}


long long prod(const int* d, int n) {
  long long v = 1;
  for (int i = 0; i < n; i++) v *= d[i];
  return v;
}

// Input (column-major) stride of every input dim.
void input_strides(const Problem& p, long long* s) {
  long long v = 1;
  for (int i = 0; i < p.rank; i++) {
    s[i] = v;
    v *= p.dim[i];
  }
}

// Output stride of every INPUT dim: input dim perm[q] sits at output
// position q, whose stride is prod_{s<q} dim[perm[s]].
void output_strides(const Problem& p, long long* s) {
  long long v = 1;
  for (int q = 0; q < p.rank; q++) {
    s[p.perm[q]] = v;
    v *= p.dim[p.perm[q]];
  }
}

int pow2ceil(int x) {
  int p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Outer map over the input dims for which `skip[d]` is false, in input order.
void build_outer(const Problem& p, const bool* skip, const long long* cin, const long long* cout,
                 int word_scale_num, int word_scale_den, OuterMap& m, int& vol_outer) {
  m.n = 0;
  vol_outer = 1;
  for (int d = 0; d < p.rank; d++) {
    if (skip[d]) continue;
    m.div[m.n] = make_fastdiv((uint32_t)p.dim[d]);
    m.ct_in[m.n] = (int)(cin[d] * word_scale_num / word_scale_den);
    m.ct_out[m.n] = (int)(cout[d] * word_scale_num / word_scale_den);
    m.n++;
    vol_outer *= p.dim[d];
  }
  for (int r = m.n; r < kMaxRank; r++) {
    m.div[r] = make_fastdiv(1);
    m.ct_in[r] = 0;
    m.ct_out[r] = 0;
  }
}

// Planning threads. Candidates of one problem are modelled in parallel by a
// small pool of worker threads that lives for the whole process (created on
// first use, never joined). One problem uses the pool at a time; a planning
// thread that finds it busy models its candidates itself. LIBRETT_PLAN_THREADS
// sets the number of threads (1: no pool).
class PlanPool {
 public:
  static PlanPool& get() {
    static PlanPool* p = new PlanPool;
    return *p;
  }
  int threads() const { return nthreads_; }
  // Runs f(0..n-1); returns false (running nothing) when the pool is busy.
  bool run(int n, const std::function<void(int)>& f) {
    std::unique_lock<std::mutex> busy(busy_mu_, std::try_to_lock);
    if (!busy.owns_lock()) return false;
    {
      std::lock_guard<std::mutex> lock(mu_);
      job_ = &f;
      n_ = n;
      next_.store(0);
      pending_ = (int)workers_.size();
      gen_++;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lock(mu_);
    done_cv_.wait(lock, [&] { return pending_ == 0; });
    job_ = nullptr;
    return true;
  }

 private:
  PlanPool() {
    const char* e = std::getenv("LIBRETT_PLAN_THREADS");
    int t = (e && *e) ? std::atoi(e) : (int)std::thread::hardware_concurrency();
    nthreads_ = std::max(1, std::min(t, 16));
    for (int i = 1; i < nthreads_; i++) {
      workers_.emplace_back([this] { loop(); });
      workers_.back().detach();
    }
  }
  void work() {
    for (;;) {
      const int i = next_.fetch_add(1);
      if (i >= n_) break;
      (*job_)(i);
    }
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      {
        std::unique_lock<std::mutex> lock(mu_);
        cv_.wait(lock, [&] { return gen_ != seen; });
        seen = gen_;
      }
      work();
      std::lock_guard<std::mutex> lock(mu_);
      if (--pending_ == 0) done_cv_.notify_all();
    }
  }
  int nthreads_ = 1;
  std::vector<std::thread> workers_;
  std::mutex busy_mu_, mu_;
  std::condition_variable cv_, done_cv_;
  const std::function<void(int)>* job_ = nullptr;
  int n_ = 0;
  std::atomic<int> next_{0};
  int pending_ = 0;
  uint64_t gen_ = 0;
};

}  // namespace

void planner_warm_up() { (void)PlanPool::get(); }

namespace {

// f(0..n-1), on the planning threads when there is enough work.
void parallel_for(int n, const std::function<void(int)>& f) {
  if (n >= 8) {
    PlanPool& pool = PlanPool::get();
    if (pool.threads() > 1 && pool.run(n, f)) return;
  }
  for (int i = 0; i < n; i++) f(i);
}

// Deterministic sample of `k` indices in [0, n).
std::vector<long long> sample_indices(long long n, int k) {
  std::vector<long long> v;
  if (n <= k) {
    for (long long i = 0; i < n; i++) v.push_back(i);
    return v;
  }
  unsigned long long s = 0x2545F4914F6CDD1DULL ^ (unsigned long long)n;
  for (int i = 0; i < k; i++) {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    v.push_back((long long)(s % (unsigned long long)n));
  }
  return v;
}

}  // namespace

void set_tile_group(TiledArgs& a, int nty, int group) {
  group = std::max(1, std::min(group, nty));
  a.group = group;
  a.group_div = make_fastdiv((uint32_t)group);
  a.group_row_div = make_fastdiv((uint32_t)(group * a.ntx));
  const int ngroups = (nty + group - 1) / group;
  const int last_rows = nty - (ngroups - 1) * group;
  a.last_div = make_fastdiv((uint32_t)last_rows);
  a.last_group_start = (ngroups - 1) * group * a.ntx;
}

Problem canonicalize(int rank, const int* dim, const int* perm, int elem) {
  Problem p;
  p.elem = elem;
  std::vector<int> newidx(rank, -1);
  int k = 0;
  long long vol = 1;
  for (int i = 0; i < rank; i++) {
    vol *= dim[i];
    if (dim[i] > 1) {
      if (k < kMaxRank) p.dim[k] = dim[i];
      newidx[i] = k++;
    }
  }
  if (k == 0) {
    p.rank = 1;
    p.dim[0] = 1;
    p.perm[0] = 0;
    p.vol = 1;
    return p;
  }
  p.rank = k;
  int j = 0;
  for (int i = 0; i < rank; i++)
    if (newidx[perm[i]] >= 0) {
      if (j < kMaxRank) p.perm[j] = newidx[perm[i]];
      j++;
    }
  p.vol = vol;
  return p;
}

Problem reduce_ranks(const Problem& p) {
  // Groups of output positions: a group continues while perm[i] == perm[i-1]+1.
  int gfirst[kMaxRank];  // first input dim of each group (output order)
  long long gvol[kMaxRank];
  int ng = 0;
  for (int i = 0; i < p.rank; i++) {
    if (i > 0 && p.perm[i] == p.perm[i - 1] + 1) {
      gvol[ng - 1] *= p.dim[p.perm[i]];
    } else {
      gfirst[ng] = p.perm[i];
      gvol[ng] = p.dim[p.perm[i]];
      ng++;
    }
  }
  // Input position of each group = rank of its first input dim.
  Problem r;
  r.elem = p.elem;
  r.rank = ng;
  r.vol = p.vol;
  for (int g = 0; g < ng; g++) {
    int pos = 0;
    for (int h = 0; h < ng; h++)
      if (gfirst[h] < gfirst[g]) pos++;
    r.perm[g] = pos;
    r.dim[pos] = (int)gvol[g];
  }
  return r;
}

void split_update(Split& s, int sizeMm, int sizeMk, const Problem& p) {
  s.sizeMm = sizeMm;
  s.sizeMk = sizeMk;
  s.volMm = (int)prod(p.dim, sizeMm);
  long long vmk = 1;
  for (int i = 0; i < sizeMk; i++) vmk *= p.dim[p.perm[i]];
  s.volMk = (int)vmk;
  bool inMmk[kMaxRank] = {false};
  for (int i = 0; i < sizeMm; i++) inMmk[i] = true;
  long long vmmk = s.volMm, vkbar = 1;
  int nmmk = sizeMm, nkbar = 0;
  for (int i = 0; i < sizeMk; i++) {
    int d = p.perm[i];
    if (d >= sizeMm) {
      vmmk *= p.dim[d];
      vkbar *= p.dim[d];
      nmmk++;
      nkbar++;
    }
    inMmk[d] = true;
  }
  s.sizeMmk = nmmk;
  s.volMmk = (int)vmmk;
  s.sizeMkBar = nkbar;
  s.volMkBar = (int)vkbar;
  s.sizeMbar = p.rank - nmmk;
  s.volMbar = (int)(p.vol / vmmk);
  if (s.splitRank >= 0) {
    s.splitDim = p.dim[s.splitRank];
    s.volMmkUnsplit = s.volMmk / s.splitDim;
  }
  const int chunk = s.splitRank >= 0 ? (s.splitDim + s.numSplit - 1) / s.numSplit : 0;
  long long cin = 1;
  for (int i = 0; i < p.rank; i++) {
    if (!inMmk[i]) break;
    if (i == s.splitRank) {
      cin *= chunk;
      break;
    }
    cin *= p.dim[i];
  }
  s.volMmkInCont = (int)cin;
  long long cout = 1;
  for (int i = 0; i < p.rank; i++) {
    int d = p.perm[i];
    if (!inMmk[d]) break;
    if (d == s.splitRank) {
      cout *= chunk;
      break;
    }
    cout *= p.dim[d];
  }
  s.volMmkOutCont = (int)cout;
}

std::string Candidate::describe() const {
  char buf[256];
  static const char* names[] = {"Unknown", "Trivial", "Packed", "PackedSplit", "Tiled", "TiledCopy"};
  const char* n = (method >= 0 && method < kNumMethods) ? names[method] : "?";
  switch (method) {
    case kTiled:
      if (xgroup > 1 || ygroup > 1)
        snprintf(buf, sizeof(buf), "%s tile %dx%d vec %d/%d groups %d/%d grid %d%s", n, tile_x, tile_y, vec_x,
                 vec_y, xgroup, ygroup, grid_x, tile_order(*this) ? " xcd" : "");
      else
        snprintf(buf, sizeof(buf), "%s tile %dx%d vec %d/%d grid %d%s", n, tile_x, tile_y, vec_x, vec_y, grid_x,
                 tile_order(*this) ? " xcd" : "");
      break;
    case kTiledCopy:
      snprintf(buf, sizeof(buf), "%s word %dB tx %d grid %d%s", n, word, copy_tx, grid_x,
               tile_order(*this) ? " xcd" : "");
      break;
    case kPacked:
    case kPackedSplit:
      snprintf(buf, sizeof(buf), "%s(%s) Mm %d Mk %d volMmk %d split %d/%d nreg %d block %d grid %dx%d lds %d%s", n,
               source, split.sizeMm, split.sizeMk, split.volMmk, split.numSplit, split.splitRank, nreg, block, grid_x,
               grid_y, lds_bytes, packed.order ? " xcd" : "");
      break;
    default:
      snprintf(buf, sizeof(buf), "%s", n);
  }
  return std::string(buf);
}

int active_blocks(const Candidate& c, const DeviceInfo& dev) {
  if (c.method == kTrivial) return 1;
  if (dev.has_gpu) {
    int n = query_occupancy(c);
    if (n > 0) return n;
  }
  const int waves = (c.block + dev.warp_size - 1) / dev.warp_size;
  int by_waves = dev.max_waves_per_cu / std::max(1, waves);
  int by_lds = c.lds_bytes > 0 ? dev.lds_per_cu / c.lds_bytes : by_waves;
  return std::max(0, std::min(by_waves, by_lds));
}

namespace {

// ---------------------------------------------------------------------------
// Tiled / TiledCopy candidates (reduced problem)
// ---------------------------------------------------------------------------

// Group map of dims `dims[0..n)` (fastest first) moving the other tensor by
// `strides`.
void make_group(const Problem& r, const int* dims, int n, const long long* strides, GroupMap& g) {
  g.n = n;
  for (int k = 0; k < kMaxGroup; k++) g.stride[k] = k < n ? (int)strides[dims[k]] : 0;
  for (int k = 0; k < kMaxGroup - 1; k++) g.div[k] = make_fastdiv(k < n ? (uint32_t)r.dim[dims[k]] : 1u);
}

// Tiled candidates of the reduced problem whose plane is x = input dims
// [0, a) and y = output dims [0, b) (the reference's Tiled is a = b = 1).
void add_tiled_plane(const Problem& r, const DeviceInfo& dev, int a_dims, int b_dims, std::vector<Candidate>& out) {
  long long cin[kMaxRank], cout[kMaxRank];
  input_strides(r, cin);
  output_strides(r, cout);
  const int B = r.elem;
  int xdims[kMaxGroup], ydims[kMaxGroup];
  long long X = 1, Y = 1;
  bool skip[kMaxRank] = {false};
  for (int k = 0; k < a_dims; k++) {
    xdims[k] = k;
    X *= r.dim[k];
    skip[k] = true;
  }
  for (int k = 0; k < b_dims; k++) {
    ydims[k] = r.perm[k];
    if (skip[ydims[k]]) return;  // the groups must not share a dim
    Y *= r.dim[ydims[k]];
  }
  for (int k = 0; k < b_dims; k++) skip[ydims[k]] = true;
  const int vmax = 16 / B;  // lanes move at most 16 bytes
  // Tile shapes (x, y). 4- and 8-byte elements also get tiles whose rows
  // are 512 bytes long in one or both directions: on MI355X they move 3-18%
  // faster than 64x64 where both dims allow 16-byte vectors (f32), less for
  // f64 (the planner's model, librettPlanMeasure or the selector decide).
  static const int shapes4[7][2] = {{64, 64}, {32, 32}, {64, 32}, {32, 64}, {128, 128}, {128, 64}, {64, 128}};
  static const int shapes8[6][2] = {{64, 64}, {32, 32}, {64, 32}, {32, 64}, {128, 64}, {64, 128}};
  static const int shapes16[4][2] = {{16, 16}, {32, 32}, {32, 16}, {16, 32}};
  const int (*shapes)[2] = (B == 16) ? shapes16 : (B == 8) ? shapes8 : shapes4;
  const int nshapes = (B == 16) ? 4 : (B == 8) ? 6 : 7;
  OuterMap outer;
  int vol_outer;
  build_outer(r, skip, cin, cout, 1, 1, outer, vol_outer);
  GroupMap xg, yg;
  make_group(r, xdims, a_dims, cout, xg);
  make_group(r, ydims, b_dims, cin, yg);
  const bool general = a_dims > 1 || b_dims > 1;
  for (int s = 0; s < nshapes; s++) {
    Candidate c;
    c.method = kTiled;
    c.elem = B;
    c.source = "reduced";
    c.tile_x = shapes[s][0];
    c.tile_y = shapes[s][1];
    c.xgroup = a_dims;
    c.ygroup = b_dims;
    // Widest lanes (vx elements per input access, vy per output access)
    // that divide the plane's extents and have a kernel instantiated;
    // ties go to the larger sum. Every input offset outside the x group is
    // a multiple of X and every output offset outside the y group a
    // multiple of Y, so X % vx == 0 and Y % vy == 0 keep lanes aligned.
    int vx = 0, vy = 0;
    for (int ax : {4, 2, 1})
      for (int ay : {4, 2, 1}) {
        if (ax > vmax || ay > vmax || X % ax || Y % ay) continue;
        if (dev.batch_mode ? !batch_tiled_available(B, c.tile_x, c.tile_y, ax, ay, general)
                           : !tiled_shape_available(B, c.tile_x, c.tile_y, ax, ay, general))
          continue;
        if (ax * ay > vx * vy || (ax * ay == vx * vy && ax + ay > vx + vy)) {
          vx = ax;
          vy = ay;
        }
      }
    if (vx == 0) continue;
    c.vec_x = vx;
    c.vec_y = vy;
    TiledArgs& a = c.tiled;
    a.X = (int)X;
    a.Y = (int)Y;
    a.in_ys = (int)cin[ydims[0]];
    a.out_xs = (int)cout[0];
    a.xg = xg;
    a.yg = yg;
    a.ntx = (int)((X + c.tile_x - 1) / c.tile_x);
    const int nty = (int)((Y + c.tile_y - 1) / c.tile_y);
    a.ntiles = a.ntx * nty;
    a.ntx_div = make_fastdiv((uint32_t)a.ntx);
    a.ntiles_div = make_fastdiv((uint32_t)a.ntiles);
    a.items = (int)std::min<long long>((long long)a.ntiles * vol_outer, 0x7fffffffLL);
    set_tile_group(a, nty, tile_group());
    a.xcds = 1;
    a.outer = outer;
    c.block = 256;
    // Group planes keep the tile's row and column offsets after the tile.
    c.lds_bytes = c.tile_y * (c.tile_x + 1) * B + (general ? ((c.tile_x + c.tile_y) * 4 + B - 1) / B * B : 0);
    c.grid_x = std::max(1, std::min(a.items, dev.cu_count * grid_per_cu()));
    c.bytes = 2 * r.vol * B;
    c.split.sizeMm = a_dims;
    c.split.sizeMk = b_dims;
    c.split.volMm = (int)X;
    c.split.volMk = (int)Y;
    c.split.volMbar = vol_outer;
    c.split.volMmk = (int)(X * Y);
    out.push_back(c);
  }
}

// Tiled candidates: the plane of input dim 0 and output dim 0 (as the
// reference), and, where those dims are short, planes whose x is the
// shortest run of leading input dims reaching 64 elements and whose y is
// the same on the output side, so that both the reads and the writes of a
// tile stay long.
void add_tiled(const Problem& r, const DeviceInfo& dev, std::vector<Candidate>& out) {
  add_tiled_plane(r, dev, 1, 1, out);
  auto reach = [&](bool input_side) {
    long long v = 1;
    int k = 0;
    while (k < r.rank && k < kMaxGroup) {
      v *= r.dim[input_side ? k : r.perm[k]];
      k++;
      if (v >= 64) break;
    }
    return k;
  };
  const int a64 = reach(true), b64 = reach(false);
  // Longest run of leading output dims that holds none of the x group's
  // input dims [0, a): with it the output rows of a tile's columns lie back
  // to back, so the tiles that run together write one compact span instead
  // of short pieces of rows far apart (4-12% faster on MI355X where the
  // output's leading dims are short; the same run on the input side
  // measured no gain).
  auto span = [&](int a) {
    long long v = 1;
    int k = 0;
    while (k < r.rank && k < kMaxGroup) {
      const int d = r.perm[k];
      if (d < a || v * r.dim[d] > (1LL << 30)) break;
      v *= r.dim[d];
      k++;
    }
    return k;
  };
  const int combos[5][2] = {{a64, 1}, {1, b64}, {a64, b64}, {1, span(1)}, {a64, span(a64)}};
  for (int i = 0; i < 5; i++) {
    const int* ab = combos[i];
    if (ab[0] < 1 || ab[1] < 1 || (ab[0] == 1 && ab[1] == 1)) continue;
    if (i < 3 && ((ab[0] > 1 && r.dim[0] >= 64) || (ab[1] > 1 && r.dim[r.perm[0]] >= 64))) continue;
    bool dup = false;
    for (int j = 0; j < i; j++) dup |= combos[j][0] == ab[0] && combos[j][1] == ab[1];
    if (!dup) add_tiled_plane(r, dev, ab[0], ab[1], out);
  }
}

void add_tiled_copy(const Problem& r, const DeviceInfo& dev, std::vector<Candidate>& out) {
  long long cin[kMaxRank], cout[kMaxRank];
  input_strides(r, cin);
  output_strides(r, cout);
  const int B = r.elem;
  const long long rowbytes = (long long)r.dim[0] * B;
  int W = 4;
  if (rowbytes % 16 == 0) W = 16;
  else if (rowbytes % 8 == 0) W = 8;
  const int X = (int)(rowbytes / W);
  const int ydim = r.perm[1];
  bool skip[kMaxRank] = {false};
  skip[0] = true;
  skip[ydim] = true;
  OuterMap outer;
  int vol_outer;
  build_outer(r, skip, cin, cout, B, W, outer, vol_outer);
  int txs[2] = {std::min(64, std::max(4, pow2ceil(X))), 64};
  if (dev.batch_mode) {
    txs[0] = X <= 16 ? 16 : 64;
    txs[1] = 64;
  }
  for (int k = 0; k < 2; k++) {
    if (k == 1 && txs[1] == txs[0]) continue;
    Candidate c;
    c.method = kTiledCopy;
    c.elem = B;
    c.source = "reduced";
    c.word = W;
    c.copy_tx = txs[k];
    if (dev.batch_mode ? !batch_copy_available(W, c.copy_tx) : !copy_shape_available(W, c.copy_tx)) continue;
    CopyArgs& a = c.copy;
    a.X = X;
    a.Y = r.dim[ydim];
    a.in_ys = (int)(cin[ydim] * B / W);
    a.out_ys = (int)(cout[ydim] * B / W);
    const int ty = (256 / c.copy_tx) * 4;
    a.ntx = (X + c.copy_tx - 1) / c.copy_tx;
    const int nty = (a.Y + ty - 1) / ty;
    a.ntiles = a.ntx * nty;
    a.ntx_div = make_fastdiv((uint32_t)a.ntx);
    a.ntiles_div = make_fastdiv((uint32_t)a.ntiles);
    a.items = (int)std::min<long long>((long long)a.ntiles * vol_outer, 0x7fffffffLL);
    a.xcds = 1;
    a.outer = outer;
    c.block = 256;
    c.lds_bytes = 0;
    c.grid_x = std::max(1, std::min(a.items, dev.cu_count * grid_per_cu()));
    c.bytes = 2 * r.vol * B;
    c.split.sizeMm = 1;
    c.split.sizeMk = 2;
    c.split.volMm = r.dim[0];
    c.split.volMkBar = a.Y;
    c.split.volMbar = vol_outer;
    out.push_back(c);
  }
}

// ---------------------------------------------------------------------------
// Packed / PackedSplit candidates
// ---------------------------------------------------------------------------

// Threads and registers for a staged volume: the smallest register count
// whose thread count fits in 512, else in the device's block limit.
bool choose_threads(int vol, int elem, const DeviceInfo& dev, int& nt, int& nreg) {
  if (dev.batch_mode) {
    // Fixed 256-thread workgroups; the smallest instantiated register count.
    for (int r = 1; r <= 16; r++) {
      if (!batch_packed_available(elem, r)) continue;
      if (256 * r >= vol) {
        nt = 256;
        nreg = r;
        return true;
      }
    }
    return false;
  }
  const int maxreg = packed_max_nreg(elem);
  for (int limit : {512, dev.max_threads_per_block}) {
    for (int r = 1; r <= maxreg; r++) {
      if (!packed_nreg_available(elem, r)) continue;
      int t = ((vol + 64 * r - 1) / (64 * r)) * 64;
      if (t <= limit) {
        nt = t;
        nreg = r;
        return true;
      }
    }
  }
  return false;
}

void fill_shape(const Problem& p, const bool* inMmk, const long long* cin, const long long* cout, int splitRank,
                int chunk, StagedShape& s) {
  int dims[kMaxRank];
  for (int d = 0; d < p.rank; d++) dims[d] = (d == splitRank) ? chunk : p.dim[d];
  long long shs[kMaxRank];
  long long v = 1;
  s.n = 0;
  for (int d = 0; d < p.rank; d++) {
    if (!inMmk[d]) continue;
    s.din[s.n] = make_fastdiv((uint32_t)std::max(1, dims[d]));
    s.ct_in[s.n] = (int)cin[d];
    shs[d] = v;
    v *= dims[d];
    s.n++;
  }
  s.vol = (int)v;
  // Pad the LDS layout when the stride of output dim 0 is even.
  const long long S = shs[p.perm[0]];
  const bool pad = S > 1 && S % 2 == 0;
  s.pad = make_fastdiv(pad ? (uint32_t)S : 0x7fffffffu);
  int k = 0;
  for (int q = 0; q < p.rank; q++) {
    int d = p.perm[q];
    if (!inMmk[d]) continue;
    s.dout[k] = make_fastdiv((uint32_t)std::max(1, dims[d]));
    s.ct_out[k] = (int)cout[d];
    s.ct_sh[k] = (int)((pad && shs[d] >= S) ? shs[d] / S * (S + 1) : shs[d]);
    k++;
  }
  for (int r = s.n; r < kMaxMmk; r++) {
    s.din[r] = make_fastdiv(1);
    s.dout[r] = make_fastdiv(1);
    s.ct_in[r] = s.ct_out[r] = s.ct_sh[r] = 0;
  }
}

// LDS elements a staged shape occupies with its padding.
int padded_elems(const StagedShape& s) {
  if (s.vol <= 0) return 1;
  return s.vol + (int)fdiv((uint32_t)(s.vol - 1), s.pad) + 1;
}

bool same_packed(const Candidate& a, const Candidate& b) {
  return a.method == b.method && a.split.volMmk == b.split.volMmk && a.split.volMbar == b.split.volMbar &&
         a.split.volMmkInCont == b.split.volMmkInCont && a.split.volMmkOutCont == b.split.volMmkOutCont &&
         a.split.numSplit == b.split.numSplit && a.split.splitRank == b.split.splitRank &&
         a.split.volMmkUnsplit == b.split.volMmkUnsplit && a.nreg == b.nreg && a.block == b.block;
}

void push_packed(const Problem& p, const char* source, const Split& s, int nt, int nreg, const DeviceInfo& dev,
                 std::vector<Candidate>& out) {
  Candidate c;
  c.method = s.numSplit > 1 || s.splitRank >= 0 ? kPackedSplit : kPacked;
  c.elem = p.elem;
  c.source = source;
  c.split = s;
  c.nreg = nreg;
  c.block = nt;
  c.bytes = 2 * p.vol * p.elem;
  for (const auto& o : out)
    if (same_packed(o, c)) return;

  long long cin[kMaxRank], cout[kMaxRank];
  input_strides(p, cin);
  output_strides(p, cout);
  bool inMmk[kMaxRank] = {false};
  for (int i = 0; i < s.sizeMm; i++) inMmk[i] = true;
  for (int i = 0; i < s.sizeMk; i++) inMmk[p.perm[i]] = true;

  PackedArgs& a = c.packed;
  if (c.method == kPacked) {
    a.num_split = 1;
    a.split_dim = 1;
    a.split_in_stride = 0;
    a.split_out_stride = 0;
    fill_shape(p, inMmk, cin, cout, -1, 0, a.shape[0]);
    a.shape[1] = a.shape[0];
    c.lds_bytes = padded_elems(a.shape[0]) * p.elem;
  } else {
    a.num_split = s.numSplit;
    a.split_dim = s.splitDim;
    a.split_in_stride = (int)cin[s.splitRank];
    a.split_out_stride = (int)cout[s.splitRank];
    const int q = s.splitDim / s.numSplit;
    fill_shape(p, inMmk, cin, cout, s.splitRank, q, a.shape[0]);
    fill_shape(p, inMmk, cin, cout, s.splitRank, q + 1, a.shape[1]);
    c.lds_bytes = std::max(padded_elems(a.shape[0]), padded_elems(a.shape[1])) * p.elem;
  }
  if (c.lds_bytes > dev.lds_per_block) return;
  if (dev.batch_mode && c.lds_bytes > kBatchLds) return;
  int vol_outer;
  build_outer(p, inMmk, cin, cout, 1, 1, a.outer, vol_outer);
  a.vol_outer = vol_outer;
  c.grid_x = a.num_split;
  int act = active_blocks(c, dev);
  if (act <= 0) return;
  c.active_blocks_per_cu = act;
  const long long want = std::max<long long>(1, (long long)dev.cu_count * act * 2 / a.num_split);
  const int gy = (int)std::max<long long>(1, std::min<long long>(vol_outer, want));
  a.outer_per_block = (vol_outer + gy - 1) / gy;
  c.grid_y = (vol_outer + a.outer_per_block - 1) / a.outer_per_block;
  a.order = 0;
  a.xcds = std::max(1, dev.cu_count / 32);  // 32 CUs per XCD on gfx950
  out.push_back(c);
  // The XCD-interleaved order of the same plan (the batched engine has its
  // own unit order).
  if (!dev.batch_mode && c.grid_y > 1 && a.outer_per_block > 1) {
    c.packed.order = 1;
    out.push_back(c);
  }
}

void add_packed(const Problem& p, const char* source, const DeviceInfo& dev, std::vector<Candidate>& out) {
  const int B = p.elem;
  // Staged bytes a workgroup may use (the batched kernel keeps room for padding).
  const long long lds_cap = dev.batch_mode ? std::min(dev.lds_per_block, kBatchLds - 1024) : dev.lds_per_block;
  // Distinct staged sets over all (Mm, Mk) splits (many pairs give the same set).
  std::vector<unsigned> seen;
  for (int numMm = 1; numMm < p.rank; numMm++) {
    for (int numMk = 1; numMk < p.rank; numMk++) {
      Split s;
      split_update(s, numMm, numMk, p);
      if (s.sizeMmk > kMaxMmk) continue;
      unsigned mask = 0;
      for (int i = 0; i < numMm; i++) mask |= 1u << i;
      for (int i = 0; i < numMk; i++) mask |= 1u << p.perm[i];
      if (std::find(seen.begin(), seen.end(), mask) != seen.end()) continue;
      seen.push_back(mask);
      const long long bytes = (long long)s.volMmk * B;
      int nt, nreg;
      if (bytes <= lds_cap && choose_threads(s.volMmk, B, dev, nt, nreg))
        push_packed(p, source, s, nt, nreg, dev, out);
      // PackedSplit: chunk one staged dim so that a stage is about 8..128 KiB.
      // Splitting a dim that starts neither contiguous run keeps both runs
      // long, so every staged dim is tried, not only the largest (the
      // reference splits the largest, src/plan.cpp:597-611).
      for (int d = 0; d < p.rank; d++) {
        if (!(mask & (1u << d)) || p.dim[d] < 2) continue;
        Split t = s;
        t.splitRank = d;
        t.numSplit = 1;
        split_update(t, numMm, numMk, p);
        const long long unsplit_bytes = (long long)t.volMmkUnsplit * B;
        int last_ns = -1;
        for (long long target : {8LL << 10, 16LL << 10, 32LL << 10, 64LL << 10, 128LL << 10, 1LL << 40}) {
          const bool last = target >= lds_cap;
          if (last) target = lds_cap;
          long long ns = (t.splitDim * unsplit_bytes + target - 1) / target;
          if (ns < 2) continue;
          if (ns > t.splitDim) ns = t.splitDim;
          const int chunk = (int)((t.splitDim + ns - 1) / ns);
          ns = (t.splitDim + chunk - 1) / chunk;  // fewest splits with this chunk size
          if (ns < 2 || ns == last_ns) continue;
          if ((long long)t.splitDim * (ns + 1) >= (1LL << 31)) continue;
          if ((long long)chunk * unsplit_bytes > lds_cap) continue;
          Split u = t;
          u.numSplit = (int)ns;
          split_update(u, numMm, numMk, p);
          if (!choose_threads(chunk * u.volMmkUnsplit, B, dev, nt, nreg)) continue;
          last_ns = (int)ns;
          push_packed(p, source, u, nt, nreg, dev, out);
          if (last) break;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// gfx950 cost model
// ---------------------------------------------------------------------------

// gfx950 cost model. A candidate's time is the launch overhead plus the
// largest of four bounds:
//   HBM      distinct 128-byte lines read and written (counted over a window
//            of consecutive rounds of one block, see model_packed) at a
//            sustained fraction of the HBM peak;
//   CU       the same bytes through the CUs that have work, each limited to
//            a per-CU bandwidth (small grids cannot fill HBM);
//   requests 128-byte transactions of all load and store wave instructions
//            (the reference's gld_tran / gst_tran, src/GpuModel.cpp:50-83)
//            at a per-CU request rate; store requests weigh more;
//   rounds   the serial load/store rounds of one resident block, each a
//            memory latency plus the round's bytes at one block's bandwidth.
// The constants were fitted (weighted toward each problem's fastest plans)
// to librettPlanMeasure-style timings of about 3200 candidate plans of 115
// problems on MI355X (55 of about 256 MiB, 60 of 1-64 MiB, 4/8/16-byte
// elements), each timed in a steady stream over buffers that do not fit the
// 256 MiB MALL, so every run pays for its own HBM reads and write-backs.
struct ModelConst {
  double bw_frac = 0.8392;         // sustained fraction of the HBM peak
  double write_weight = 1.4047;    // written line vs read line
  double partial_weight = 0.0;     // extra weight of a partially written line
  double round_latency = 1.325e-6; // latency part of one block round (s)
  double launch = 3.998e-6;        // launch and completion overhead (s)
  double tran_time = 0.1974e-9;    // seconds per load transaction per CU
  double store_tran_weight = 63.65;// store transaction vs load transaction
  double cu_bw = 27.39e9;          // bytes/s one CU can move
  double block_bw = 15.06e9;       // bytes/s of one block within its rounds
};

ModelConst g_model;

struct Run {
  long long s, n;  // byte offset and length
};

// Distinct 128-byte lines a set of runs touches, and how many of them the
// runs cover only in part (a run boundary falls inside the line). `runs`
// must be sorted by start offset.
void line_stats_sorted(const std::vector<Run>& runs, double& lines, double& partial) {
  lines = partial = 0;
  if (runs.empty()) return;
  long long last_line = -1, last_partial = -1;
  size_t i = 0;
  while (i < runs.size()) {
    long long s0 = runs[i].s, e0 = runs[i].s + runs[i].n;
    size_t j = i + 1;
    while (j < runs.size() && runs[j].s <= e0) {
      e0 = std::max(e0, runs[j].s + runs[j].n);
      j++;
    }
    const long long first = s0 >> 7, last = (e0 - 1) >> 7;
    lines += (double)(last - first + 1) - (first == last_line ? 1 : 0);
    if ((s0 & 127) && first != last_partial) {
      partial += 1;
      last_partial = first;
    }
    if ((e0 & 127) && last != last_partial) {
      partial += 1;
      last_partial = last;
    }
    last_line = last;
    i = j;
  }
}

void line_stats(std::vector<Run>& runs, double& lines, double& partial) {
  auto less = [](const Run& a, const Run& b) { return a.s < b.s; };
  if (!std::is_sorted(runs.begin(), runs.end(), less)) std::sort(runs.begin(), runs.end(), less);
  line_stats_sorted(runs, lines, partial);
}

// Runs merged the way line_stats_sorted merges them (a run that starts at
// or before the current end joins it). `runs` sorted by start.
void merged_intervals(const std::vector<Run>& runs, std::vector<Run>& out) {
  out.clear();
  for (const Run& r : runs) {
    if (!out.empty() && r.s <= out.back().s + out.back().n)
      out.back().n = std::max(out.back().s + out.back().n, r.s + r.n) - out.back().s;
    else
      out.push_back(r);
  }
}

// Whether line_stats counts line L as partly covered by `runs` (any order):
// a merged interval of them starts or ends inside L, off its boundary.
bool line_partial(long long L, std::vector<Run> runs) {
  std::sort(runs.begin(), runs.end(), [](const Run& a, const Run& b) { return a.s < b.s; });
  std::vector<Run> m;
  merged_intervals(runs, m);
  for (const Run& r : m) {
    const long long e = r.s + r.n;
    if (((r.s >> 7) == L && (r.s & 127)) || (((e - 1) >> 7) == L && (e & 127))) return true;
  }
  return false;
}

// Merges consecutive sorted segments of `runs` (segment k ends at ends[k])
// into one sorted list: the exact order line_stats would sort them into,
// without a full sort.
void merge_segments(std::vector<Run>& runs, std::vector<size_t>& ends) {
  auto less = [](const Run& a, const Run& b) { return a.s < b.s; };
  if (std::is_sorted(runs.begin(), runs.end(), less)) return;
  while (ends.size() > 1) {
    std::vector<size_t> next;
    size_t begin = 0;
    for (size_t k = 0; k < ends.size(); k += 2) {
      if (k + 1 < ends.size()) {
        std::inplace_merge(runs.begin() + begin, runs.begin() + ends[k], runs.begin() + ends[k + 1], less);
        next.push_back(ends[k + 1]);
        begin = ends[k + 1];
      } else {
        next.push_back(ends[k]);
      }
    }
    ends.swap(next);
  }
}

// 128-byte transactions of one run of `n` bytes at byte offset `s` (one
// wave instruction touching it issues one request per line it spans).
inline double tran(long long s, long long n) { return n <= 0 ? 0 : (double)(((s + n - 1) >> 7) - (s >> 7) + 1); }

// Transactions of a staged shape's runs (element offsets) for one round:
// runs are walked in LDS order (input side) or output order, cut at 64-lane
// wave boundaries.
double staged_tran(const std::vector<std::pair<long long, int>>& runs, long long base, int B) {
  double t = 0;
  long long pos = 0;  // position in walk order
  for (const auto& r : runs) {
    long long s = r.first, n = r.second;
    while (n > 0) {
      const long long room = 64 - (pos & 63);
      const long long k = std::min<long long>(room, n);
      t += tran((base + s) * B, k * B);
      s += k;
      n -= k;
      pos += k;
    }
  }
  return t;
}

void model_tiled(Candidate& c) {
  const TiledArgs& a = c.tiled;
  const int B = c.elem;
  double R = 0, W = 0, P = 0, TR = 0, TW = 0;
  std::vector<long long> ws = sample_indices(a.items, 24);
  std::vector<Run> rin, rout;
  for (long long w : ws) {
    uint32_t mb = (uint32_t)(w / a.ntiles);
    int tile = (int)(w % a.ntiles);
    int x0 = (tile % a.ntx) * c.tile_x, y0 = (tile / a.ntx) * c.tile_y;
    int oi, oo;
    outer_offsets(a.outer, mb, oi, oo);
    int tx = std::min(c.tile_x, a.X - x0), ty = std::min(c.tile_y, a.Y - y0);
    rin.clear();
    rout.clear();
    for (int y = 0; y < ty; y++)
      rin.push_back({(oi + x0 + (long long)tiled_in_y(a, (uint32_t)(y0 + y))) * B, (long long)tx * B});
    for (int x = 0; x < tx; x++)
      rout.push_back({(oo + y0 + (long long)tiled_out_x(a, (uint32_t)(x0 + x))) * B, (long long)ty * B});
    for (const Run& r : rin) TR += tran(r.s, r.n);
    for (const Run& r : rout) TW += tran(r.s, r.n);
    double l, p;
    line_stats(rin, l, p);
    R += l;
    line_stats(rout, l, p);
    W += l;
    P += p;
  }
  const double scale = ws.empty() ? 0 : (double)a.items / ws.size();
  c.model.read_lines = R * scale;
  c.model.write_lines = W * scale;
  c.model.partial_lines = P * scale;
  c.model.read_tran = TR * scale;
  c.model.write_tran = TW * scale;
  c.model.items = a.items;
  c.model.steps_per_block = (double)a.items / c.grid_x;
}

void model_copy(Candidate& c) {
  const CopyArgs& a = c.copy;
  const int Wb = c.word;
  const int ty_tile = (256 / c.copy_tx) * 4;
  double R = 0, W = 0, P = 0, TR = 0, TW = 0;
  std::vector<long long> ws = sample_indices(a.items, 24);
  std::vector<Run> rin, rout;
  for (long long w : ws) {
    uint32_t mb = (uint32_t)(w / a.ntiles);
    int tile = (int)(w % a.ntiles);
    int x0 = (tile % a.ntx) * c.copy_tx, y0 = (tile / a.ntx) * ty_tile;
    int oi, oo;
    outer_offsets(a.outer, mb, oi, oo);
    int tx = std::min(c.copy_tx, a.X - x0), ty = std::min(ty_tile, a.Y - y0);
    rin.clear();
    rout.clear();
    for (int y = 0; y < ty; y++) {
      rin.push_back({(oi + x0 + (long long)(y0 + y) * a.in_ys) * Wb, (long long)tx * Wb});
      rout.push_back({(oo + x0 + (long long)(y0 + y) * a.out_ys) * Wb, (long long)tx * Wb});
    }
    for (const Run& r : rin) TR += tran(r.s, r.n);
    for (const Run& r : rout) TW += tran(r.s, r.n);
    double l, p;
    line_stats(rin, l, p);
    R += l;
    line_stats(rout, l, p);
    W += l;
    P += p;
  }
  const double scale = ws.empty() ? 0 : (double)a.items / ws.size();
  c.model.read_lines = R * scale;
  c.model.write_lines = W * scale;
  c.model.partial_lines = P * scale;
  c.model.read_tran = TR * scale;
  c.model.write_tran = TW * scale;
  c.model.items = a.items;
  c.model.steps_per_block = (double)a.items / c.grid_x;
}

}  // namespace

void staged_runs(const StagedShape& s, bool output_side, std::vector<std::pair<long long, int>>& runs) {
  runs.clear();
  const int n = s.n;
  int dd[kMaxMmk];
  long long ct[kMaxMmk];
  for (int d = 0; d < n; d++) {
    dd[d] = output_side ? (int)s.dout[d].d : (int)s.din[d].d;
    ct[d] = output_side ? s.ct_out[d] : s.ct_in[d];
  }
  // A staged shape is a sub-box of a column-major array walked in that
  // array's own order, so consecutive elements are adjacent exactly within
  // the leading dims whose strides make them contiguous (dims of extent 1
  // take no step): each run is one block of those dims. The blocks are
  // listed by an odometer over the remaining dims, one step per run.
  long long len = 1;
  int k = 0;
  while (k < n && (dd[k] == 1 || ct[k] == len)) {
    len *= dd[k];
    k++;
  }
  const long long nruns = len > 0 ? (long long)s.vol / len : 0;
  runs.reserve((size_t)nruns);
  int dig[kMaxMmk] = {0};
  long long off = 0;
  for (long long r = 0; r < nruns; r++) {
    runs.push_back({off, (int)len});
    for (int d = k; d < n; d++) {
      if (++dig[d] < dd[d]) {
        off += ct[d];
        break;
      }
      off -= (long long)(dd[d] - 1) * ct[d];
      dig[d] = 0;
    }
  }
}

namespace {

// Packed / PackedSplit: lines are counted over a window of consecutive outer
// indices of one block, because a block walks its outer range in order and
// the L2 keeps the lines its previous rounds touched; short runs that are
// adjacent across rounds therefore cost one line fetch, not one per run.
//
// The sums depend only on the candidate's staged shapes, split and outer
// map, not on its work order, so the XCD-interleaved twin of a plan reuses
// them (PackedSums). One round's own line and transaction counts depend only
// on where its base offset falls within a 128-byte line, so they are
// computed once per (chunk variant, side, alignment). A staged shape walked
// in its own order has strictly increasing offsets (a sub-box of a
// column-major array), so each round's runs come sorted and a window needs a
// merge, not a sort.
struct PackedSums {
  double R = 0, W = 0, P = 0, TR = 0, TW = 0, rounds = 0, R1 = 0, W1 = 0;
};

// Statistics of one round's runs placed at byte alignment `al` within a
// line: distinct lines, partly covered lines, transactions, the first and
// last line touched (relative to the line holding the base) and the merged
// intervals that touch those two lines.
struct RoundPattern {
  bool done = false;
  double lines = 0, partial = 0, tran = 0;
  long long first = 0, last = -1;
  std::vector<Run> head, tail;
};

// Lines and partly covered lines of the union of several rounds of one
// pattern side, from each round's own statistics. Rounds whose line ranges
// are disjoint, or share only a boundary line (with no line shared by three
// rounds), combine exactly: shared lines are counted once and their partial
// state is recomputed from the intervals that touch them. Returns false
// when the ranges overlap further (the caller then merges the runs).
struct PlacedRound {
  long long q;  // line of the round's base
  const RoundPattern* pat;
};
bool union_stats(std::vector<PlacedRound>& rs, double& lines, double& partial) {
  std::sort(rs.begin(), rs.end(), [](const PlacedRound& a, const PlacedRound& b) {
    return a.q + a.pat->first < b.q + b.pat->first;
  });
  lines = partial = 0;
  for (size_t k = 0; k < rs.size(); k++) {
    lines += rs[k].pat->lines;
    partial += rs[k].pat->partial;
    if (k == 0) continue;
    const PlacedRound& A = rs[k - 1];
    const PlacedRound& Bq = rs[k];
    const long long a_last = A.q + A.pat->last, b_first = Bq.q + Bq.pat->first;
    if (b_first > a_last) continue;
    if (b_first < a_last) return false;
    // Shared line L: a line shared with the round before A as well is left
    // to the caller.
    if (k >= 2 && A.pat->first == A.pat->last && rs[k - 2].q + rs[k - 2].pat->last == a_last) return false;
    if (Bq.pat->first == Bq.pat->last && k + 1 < rs.size() && rs[k + 1].q + rs[k + 1].pat->first == a_last)
      return false;
    const long long L = a_last;
    std::vector<Run> ta, hb, both;
    for (const Run& r : A.pat->tail) ta.push_back({r.s + A.q * 128, r.n});
    for (const Run& r : Bq.pat->head) hb.push_back({r.s + Bq.q * 128, r.n});
    both = ta;
    both.insert(both.end(), hb.begin(), hb.end());
    lines -= 1;
    partial += (line_partial(L, both) ? 1 : 0) - (line_partial(L, ta) ? 1 : 0) - (line_partial(L, hb) ? 1 : 0);
  }
  return true;
}

PackedSums packed_sums(const Candidate& c) {
  const PackedArgs& a = c.packed;
  const int B = c.elem;
  PackedSums S;
  std::vector<std::pair<long long, int>> rin[2], rout[2];
  const int nvar = (c.method == kPackedSplit) ? 2 : 1;
  for (int v = 0; v < nvar; v++) {
    staged_runs(a.shape[v], false, rin[v]);
    staged_runs(a.shape[v], true, rout[v]);
  }
  // Round patterns of variant v, side (0 in, 1 out) at byte alignment `al`
  // = (base * B) mod 128.
  std::vector<RoundPattern> memo(2 * 2 * 128);
  std::vector<Run> tmp, merged;
  auto pattern = [&](int v, int side, long long al) -> const RoundPattern& {
    RoundPattern& r = memo[(v * 2 + side) * 128 + al];
    if (!r.done) {
      const auto& runs = side ? rout[v] : rin[v];
      tmp.clear();
      for (const auto& x : runs) tmp.push_back({al + x.first * B, (long long)x.second * B});
      line_stats_sorted(tmp, r.lines, r.partial);
      r.tran = staged_tran(runs, al / B, B);
      merged_intervals(tmp, merged);
      if (!merged.empty()) {
        r.first = merged.front().s >> 7;
        r.last = (merged.back().s + merged.back().n - 1) >> 7;
        for (const Run& m : merged) {
          if ((m.s >> 7) <= r.first) r.head.push_back(m);
          if (((m.s + m.n - 1) >> 7) >= r.last) r.tail.push_back(m);
        }
      }
      r.done = true;
    }
    return r;
  };
  const long long nblocks = (long long)c.grid_x * c.grid_y;
  const int window = std::max(1, std::min(4, a.outer_per_block));
  std::vector<long long> bs = sample_indices(nblocks, 6);
  std::vector<Run> win, wout;
  std::vector<size_t> ein, eout;
  std::vector<PlacedRound> pin, pout;
  for (long long b : bs) {
    const int split = (int)(b % c.grid_x);
    const int by = (int)(b / c.grid_x);
    const int q0 = a.split_dim / a.num_split;
    const int c0 = (int)((long long)split * a.split_dim / a.num_split);
    const int c1 = (int)((long long)(split + 1) * a.split_dim / a.num_split);
    const int plus = (nvar == 2) ? (c1 - c0) - q0 : 0;
    const int o0 = by * a.outer_per_block;
    const int o1 = std::min(a.vol_outer, o0 + window);
    pin.clear();
    pout.clear();
    for (int o = o0; o < o1; o++) {
      int oi, oo;
      outer_offsets(a.outer, (uint32_t)o, oi, oo);
      const long long bi = ((long long)c0 * a.split_in_stride + oi) * B;
      const long long bo = ((long long)c0 * a.split_out_stride + oo) * B;
      // Each round alone (no reuse between rounds).
      const RoundPattern& ri = pattern(plus, 0, bi & 127);
      const RoundPattern& ro = pattern(plus, 1, bo & 127);
      S.TR += ri.tran;
      S.TW += ro.tran;
      S.R1 += ri.lines;
      S.W1 += ro.lines;
      pin.push_back({bi >> 7, &ri});
      pout.push_back({bo >> 7, &ro});
    }
    double l, p;
    if (!union_stats(pin, l, p)) {
      win.clear();
      ein.clear();
      for (int o = o0; o < o1; o++) {
        int oi, oo;
        outer_offsets(a.outer, (uint32_t)o, oi, oo);
        const long long bi = (long long)c0 * a.split_in_stride + oi;
        for (const auto& r : rin[plus]) win.push_back({(bi + r.first) * B, (long long)r.second * B});
        ein.push_back(win.size());
      }
      merge_segments(win, ein);
      line_stats_sorted(win, l, p);
    }
    S.R += l;
    if (!union_stats(pout, l, p)) {
      wout.clear();
      eout.clear();
      for (int o = o0; o < o1; o++) {
        int oi, oo;
        outer_offsets(a.outer, (uint32_t)o, oi, oo);
        const long long bo = (long long)c0 * a.split_out_stride + oo;
        for (const auto& r : rout[plus]) wout.push_back({(bo + r.first) * B, (long long)r.second * B});
        eout.push_back(wout.size());
      }
      merge_segments(wout, eout);
      line_stats_sorted(wout, l, p);
    }
    S.W += l;
    S.P += p;
    S.rounds += o1 - o0;
  }
  return S;
}

void model_packed(Candidate& c, const DeviceInfo& dev, const PackedSums& S) {
  const PackedArgs& a = c.packed;
  const long long nitems = (long long)a.num_split * a.vol_outer;
  const double scale = S.rounds > 0 ? (double)nitems / S.rounds : 0;
  // Reuse between rounds needs the lines of one round of all resident blocks
  // to stay in L2; past that footprint it fades out.
  const double resident = std::min((double)c.grid_x * c.grid_y, (double)dev.cu_count * std::max(1, c.active_blocks_per_cu));
  const double footprint = resident * 128.0 * (S.R1 + S.W1) / std::max(1.0, S.rounds);
  // In the XCD-interleaved order the blocks sharing lines run together on
  // one XCD, so the reuse does not depend on the lines surviving in L2.
  const double keep = a.order == 1 ? 1.0 : std::min(1.0, dev.l2_bytes / std::max(1.0, footprint));
  c.model.read_lines = (S.R * keep + S.R1 * (1 - keep)) * scale;
  c.model.write_lines = (S.W * keep + S.W1 * (1 - keep)) * scale;
  c.model.partial_lines = S.P * scale;
  c.model.read_tran = S.TR * scale;
  c.model.write_tran = S.TW * scale;
  c.model.items = (double)nitems;
  c.model.steps_per_block = a.outer_per_block;
}

// True when b is a's plan in the other work order (same sums).
bool order_twins(const Candidate& a, const Candidate& b) {
  if (a.method == kTiled || a.method == kTiledCopy)
    return b.method == a.method && tile_order(a) != tile_order(b) && a.tile_x == b.tile_x &&
           a.tile_y == b.tile_y && a.vec_x == b.vec_x && a.vec_y == b.vec_y && a.word == b.word &&
           a.copy_tx == b.copy_tx && a.xgroup == b.xgroup && a.ygroup == b.ygroup;
  return (a.method == kPacked || a.method == kPackedSplit) && same_packed(a, b) && a.grid_x == b.grid_x &&
         a.grid_y == b.grid_y && a.packed.outer_per_block == b.packed.outer_per_block &&
         a.source == b.source && a.packed.order != b.packed.order;
}

// Every Tiled and TiledCopy plan in both tile orders across XCDs, the two
// next to each other: dispatch order (workgroup b starts at tile b) and
// XCD-contiguous order (each XCD walks one contiguous range of tiles, so
// tiles sharing 128-byte lines meet in one L2). Which is faster depends on
// the shape (measured on MI355X: +5-15% for Tiled planes whose rows are not
// line-aligned, -8% on some small group planes), so both are candidates.
// The analytic model cannot tell them apart; its tie goes to the first,
// XCD-contiguous for Tiled and dispatch order for TiledCopy, the better
// default of each on the seeded ~256 MiB cases.
void add_order_twins(const DeviceInfo& dev, std::vector<Candidate>& out) {
  const int xcds = std::max(1, dev.cu_count / 32);  // 32 CUs per XCD on gfx950
  if (xcds < 2) return;
  std::vector<Candidate> res;
  res.reserve(out.size() * 2);
  for (const Candidate& c : out) {
    if (c.method != kTiled && c.method != kTiledCopy) {
      res.push_back(c);
      continue;
    }
    Candidate x = c;
    if (x.method == kTiled)
      x.tiled.xcds = xcds;
    else
      x.copy.xcds = xcds;
    if (c.method == kTiled) {
      res.push_back(x);
      res.push_back(c);
    } else {
      res.push_back(c);
      res.push_back(x);
    }
  }
  out.swap(res);
}

}  // namespace

void set_model_constants(const double* v, int n) {
  double* f[] = {&g_model.bw_frac,   &g_model.write_weight,      &g_model.partial_weight,
                 &g_model.round_latency, &g_model.launch,        &g_model.tran_time,
                 &g_model.store_tran_weight, &g_model.cu_bw,     &g_model.block_bw};
  for (int i = 0; i < n && i < 9; i++)
    if (v[i] >= 0) *f[i] = v[i];
}

void model_candidate(Candidate& c, const Problem& full, const DeviceInfo& dev) {
  ModelStats& m = c.model;
  m = ModelStats();
  const ModelConst& K = g_model;
  if (c.method == kTrivial) {
    m.read_lines = m.write_lines = (double)(c.bytes / 2) / 128.0;
    m.t_hbm = (double)c.bytes / (dev.hbm_bytes_per_s * K.bw_frac);
    m.t_stream = m.t_hbm;
    m.t_launch = K.launch;
    m.items = 1;
    m.predicted_s = m.t_hbm + m.t_launch;
    return;
  }
  if (c.active_blocks_per_cu <= 0) c.active_blocks_per_cu = active_blocks(c, dev);
  switch (c.method) {
    case kTiled:
      model_tiled(c);
      break;
    case kTiledCopy:
      model_copy(c);
      break;
    default:
      model_packed(c, dev, packed_sums(c));
      break;
  }
  model_times(c, dev);
  (void)full;
}

void model_times(Candidate& c, const DeviceInfo& dev) {
  ModelStats& m = c.model;
  const ModelConst& K = g_model;
  if (c.method == kTrivial) return;
  const double blocks = (double)c.grid_x * c.grid_y;
  m.concurrency = std::min(blocks, (double)dev.cu_count * std::max(1, c.active_blocks_per_cu));
  const double line_bytes =
      128.0 * (m.read_lines + K.write_weight * m.write_lines + K.partial_weight * m.partial_lines);
  const double cus = std::min((double)dev.cu_count, m.concurrency);
  const double t_hbm = line_bytes / (dev.hbm_bytes_per_s * K.bw_frac);
  const double t_cu = line_bytes / (std::max(1.0, cus) * K.cu_bw);
  const double t_req = (m.read_tran + K.store_tran_weight * m.write_tran) * K.tran_time / dev.cu_count;
  const double round_bytes = (double)c.bytes / std::max(1.0, m.items);
  const double rounds = std::ceil(m.items / std::max(1.0, m.concurrency));
  m.t_hbm = std::max(std::max(t_hbm, t_cu), t_req);
  m.t_stream = std::max(t_hbm, t_req);
  m.t_latency = rounds * (K.round_latency + round_bytes / K.block_bw);
  m.t_launch = K.launch;
  m.predicted_s = std::max(m.t_hbm, m.t_latency) + m.t_launch;
}

std::vector<Candidate> make_candidates(const Problem& full, const DeviceInfo& dev) {
  std::vector<Candidate> out;
  if (full.rank < 1 || full.rank > kMaxRank) return out;
  const Problem red = reduce_ranks(full);
  if (red.rank == 1) {
    Candidate c;
    c.method = kTrivial;
    c.elem = full.elem;
    c.source = "reduced";
    c.bytes = 2 * full.vol * full.elem;
    c.split.volMmk = (int)full.vol;
    out.push_back(c);
    model_candidate(out.back(), full, dev);
    return out;
  }
  if (red.perm[0] == 0)
    add_tiled_copy(red, dev, out);
  else
    add_tiled(red, dev, out);
  if (!dev.batch_mode) add_order_twins(dev, out);
  add_packed(full, "full", dev, out);
  if (red.rank != full.rank) add_packed(red, "reduced", dev, out);
  for (auto& c : out)
    if (c.active_blocks_per_cu <= 0) c.active_blocks_per_cu = active_blocks(c, dev);
  // Model every candidate; the XCD-interleaved twin of a Packed plan follows
  // it in the list and shares its sums. Candidates are independent, so the
  // work is spread over the planning threads.
  const size_t n = out.size();
  std::vector<PackedSums> sums(n);
  std::vector<char> twin(n, 0);
  for (size_t i = 1; i < n; i++) twin[i] = order_twins(out[i - 1], out[i]);
  parallel_for((int)n, [&](int i) {
    Candidate& c = out[i];
    if (c.method == kPacked || c.method == kPackedSplit) {
      if (!twin[i]) sums[i] = packed_sums(c);
    } else if (!twin[i]) {
      model_candidate(c, full, dev);
    }
  });
  // A Tiled / TiledCopy twin has the same counts (the model does not see
  // which XCD runs a tile).
  for (size_t i = 0; i < n; i++)
    if (twin[i] && (out[i].method == kTiled || out[i].method == kTiledCopy)) out[i].model = out[i - 1].model;
  for (size_t i = 0; i < n; i++) {
    Candidate& c = out[i];
    if (c.method != kPacked && c.method != kPackedSplit) continue;
    if (twin[i]) sums[i] = sums[i - 1];
    c.model = ModelStats();
    model_packed(c, dev, sums[i]);
    model_times(c, dev);
  }
  // Drop candidates that cannot launch on this device.
  std::vector<Candidate> ok;
  for (auto& c : out)
    if (c.method == kTrivial || c.active_blocks_per_cu > 0) ok.push_back(c);
  return ok;
}

int choose_by_model(const std::vector<Candidate>& cands) {
  int best = -1;
  for (int i = 0; i < (int)cands.size(); i++) {
    if (cands[i].method == kTrivial) return i;
    if (best < 0 || cands[i].model.predicted_s < cands[best].model.predicted_s) best = i;
  }
  return best;
}

Candidate scalar_fallback(const Candidate& c) {
  Candidate f = c;
  if (c.method == kTiled) {
    f.vec_x = 1;
    f.vec_y = 1;
  }
  return f;
}

int needed_alignment(const Candidate& c) {
  switch (c.method) {
    case kTiled:
      return c.elem * std::max(c.vec_x, c.vec_y);
    case kTiledCopy:
      return c.word;
    default:
      return c.elem;
  }
}

Candidate copy_fallback(const Candidate& c) {
  // TiledCopy moves W-byte words; for pointers aligned only to the element,
  // rebuild the arguments with W = element size.
  Candidate f = c;
  if (c.method != kTiledCopy || c.word == c.elem) return f;
  const int k = c.word / c.elem;
  CopyArgs& a = f.copy;
  a.X *= k;
  a.in_ys *= k;
  a.out_ys *= k;
  for (int r = 0; r < a.outer.n; r++) {
    a.outer.ct_in[r] *= k;
    a.outer.ct_out[r] *= k;
  }
  f.word = c.elem;
  f.copy_tx = 64;
  const int ty = (256 / f.copy_tx) * 4;
  a.ntx = (a.X + f.copy_tx - 1) / f.copy_tx;
  const int nty = (a.Y + ty - 1) / ty;
  const long long vol_outer = c.copy.items / std::max(1, c.copy.ntiles);
  a.ntiles = a.ntx * nty;
  a.ntx_div = make_fastdiv((uint32_t)a.ntx);
  a.ntiles_div = make_fastdiv((uint32_t)a.ntiles);
  a.items = (int)(a.ntiles * vol_outer);
  f.grid_x = std::max(1, std::min(a.items, 16384));
  return f;
}

Candidate unaligned_fallback(const Candidate& c) {
  if (c.method == kTiled) return scalar_fallback(c);
  if (c.method == kTiledCopy) return copy_fallback(c);
  return c;
}

}  // namespace lrt
