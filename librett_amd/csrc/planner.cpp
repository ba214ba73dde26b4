// Host planner for the gfx950 transpose library. See planner.h.
#include "planner.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "kernels.h"

namespace lrt {

namespace {

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

// 64-byte HBM sectors touched by a contiguous run of `len` elements of `b`
// bytes starting at element `start` (buffers are at least 256-byte aligned).
inline long long sectors(long long start, long long len, int b) {
  if (len <= 0) return 0;
  long long first = start * b / 64;
  long long last = ((start + len) * b - 1) / 64;
  return last - first + 1;
}

// Sectors of a run that the run covers only in part (first/last sector).
inline int partial_sectors(long long start, long long len, int b) {
  if (len <= 0) return 0;
  long long s0 = start * b, s1 = (start + len) * b;
  long long first = s0 / 64, last = (s1 - 1) / 64;
  int n = 0;
  if (s0 % 64 != 0) n++;
  if (s1 % 64 != 0 && (last != first || s0 % 64 == 0)) n++;
  return n;
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
      snprintf(buf, sizeof(buf), "%s tile %dx%d vec %d/%d grid %d", n, tile_x, tile_y, vec_x, vec_y, grid_x);
      break;
    case kTiledCopy:
      snprintf(buf, sizeof(buf), "%s word %dB tx %d grid %d", n, word, copy_tx, grid_x);
      break;
    case kPacked:
    case kPackedSplit:
      snprintf(buf, sizeof(buf), "%s(%s) Mm %d Mk %d volMmk %d split %d/%d nreg %d block %d grid %dx%d lds %d", n,
               source, split.sizeMm, split.sizeMk, split.volMmk, split.numSplit, split.splitRank, nreg, block, grid_x,
               grid_y, lds_bytes);
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

void add_tiled(const Problem& r, const DeviceInfo& dev, std::vector<Candidate>& out) {
  long long cin[kMaxRank], cout[kMaxRank];
  input_strides(r, cin);
  output_strides(r, cout);
  const int B = r.elem;
  const int ydim = r.perm[0];
  const int X = r.dim[0], Y = r.dim[ydim];
  const int v = 16 / B;
  const int vx = (v > 1 && X % v == 0) ? v : 1;
  const int vy = (v > 1 && Y % v == 0) ? v : 1;
  static const int shapes4[4][2] = {{64, 64}, {32, 32}, {64, 32}, {32, 64}};
  static const int shapes16[4][2] = {{16, 16}, {32, 32}, {32, 16}, {16, 32}};
  const int (*shapes)[2] = (B == 16) ? shapes16 : shapes4;
  bool skip[kMaxRank] = {false};
  skip[0] = true;
  skip[ydim] = true;
  OuterMap outer;
  int vol_outer;
  build_outer(r, skip, cin, cout, 1, 1, outer, vol_outer);
  for (int s = 0; s < 4; s++) {
    Candidate c;
    c.method = kTiled;
    c.elem = B;
    c.source = "reduced";
    c.tile_x = shapes[s][0];
    c.tile_y = shapes[s][1];
    c.vec_x = vx;
    c.vec_y = vy;
    if (!tiled_shape_available(B, c.tile_x, c.tile_y, vx, vy)) continue;
    // Skip tiles more than twice the plane in a direction when a smaller one exists.
    TiledArgs& a = c.tiled;
    a.X = X;
    a.Y = Y;
    a.in_ys = (int)cin[ydim];
    a.out_xs = (int)cout[0];
    a.ntx = (X + c.tile_x - 1) / c.tile_x;
    const int nty = (Y + c.tile_y - 1) / c.tile_y;
    a.ntiles = a.ntx * nty;
    a.ntx_div = make_fastdiv((uint32_t)a.ntx);
    a.ntiles_div = make_fastdiv((uint32_t)a.ntiles);
    a.items = (int)std::min<long long>((long long)a.ntiles * vol_outer, 0x7fffffffLL);
    a.outer = outer;
    c.block = 256;
    c.lds_bytes = c.tile_y * (c.tile_x + 1) * B;
    c.grid_x = std::max(1, std::min(a.items, dev.cu_count * 64));
    c.bytes = 2 * r.vol * B;
    c.split.sizeMm = 1;
    c.split.sizeMk = 1;
    c.split.volMm = X;
    c.split.volMk = Y;
    c.split.volMbar = vol_outer;
    c.split.volMmk = X * Y;
    out.push_back(c);
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
  for (int k = 0; k < 2; k++) {
    if (k == 1 && txs[1] == txs[0]) continue;
    Candidate c;
    c.method = kTiledCopy;
    c.elem = B;
    c.source = "reduced";
    c.word = W;
    c.copy_tx = txs[k];
    if (!copy_shape_available(W, c.copy_tx)) continue;
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
    a.outer = outer;
    c.block = 256;
    c.lds_bytes = 0;
    c.grid_x = std::max(1, std::min(a.items, dev.cu_count * 64));
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
  int k = 0;
  for (int q = 0; q < p.rank; q++) {
    int d = p.perm[q];
    if (!inMmk[d]) continue;
    s.dout[k] = make_fastdiv((uint32_t)std::max(1, dims[d]));
    s.ct_out[k] = (int)cout[d];
    s.ct_sh[k] = (int)shs[d];
    k++;
  }
  for (int r = s.n; r < kMaxMmk; r++) {
    s.din[r] = make_fastdiv(1);
    s.dout[r] = make_fastdiv(1);
    s.ct_in[r] = s.ct_out[r] = s.ct_sh[r] = 0;
  }
}

bool same_packed(const Candidate& a, const Candidate& b) {
  return a.method == b.method && a.split.volMmk == b.split.volMmk && a.split.volMbar == b.split.volMbar &&
         a.split.volMmkInCont == b.split.volMmkInCont && a.split.volMmkOutCont == b.split.volMmkOutCont &&
         a.split.numSplit == b.split.numSplit && a.nreg == b.nreg && a.block == b.block;
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
    c.lds_bytes = a.shape[0].vol * p.elem;
  } else {
    a.num_split = s.numSplit;
    a.split_dim = s.splitDim;
    a.split_in_stride = (int)cin[s.splitRank];
    a.split_out_stride = (int)cout[s.splitRank];
    const int q = s.splitDim / s.numSplit;
    fill_shape(p, inMmk, cin, cout, s.splitRank, q, a.shape[0]);
    fill_shape(p, inMmk, cin, cout, s.splitRank, q + 1, a.shape[1]);
    const int maxchunk = (s.splitDim + s.numSplit - 1) / s.numSplit;
    c.lds_bytes = maxchunk * s.volMmkUnsplit * p.elem;
  }
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
  out.push_back(c);
}

void add_packed(const Problem& p, const char* source, const DeviceInfo& dev, std::vector<Candidate>& out) {
  const int B = p.elem;
  for (int numMm = 1; numMm < p.rank; numMm++) {
    for (int numMk = 1; numMk < p.rank; numMk++) {
      Split s;
      split_update(s, numMm, numMk, p);
      if (s.sizeMmk > kMaxMmk) continue;
      const long long bytes = (long long)s.volMmk * B;
      int nt, nreg;
      if (bytes <= dev.lds_per_block && choose_threads(s.volMmk, B, dev, nt, nreg)) {
        push_packed(p, source, s, nt, nreg, dev, out);
        continue;
      }
      // Split the largest staged dim (first in input order, then output order).
      Split t = s;
      int maxDim = 0;
      for (int i = 0; i < t.sizeMm; i++)
        if (p.dim[i] > maxDim) {
          maxDim = p.dim[i];
          t.splitRank = i;
        }
      for (int i = 0; i < t.sizeMk; i++) {
        int d = p.perm[i];
        if (p.dim[d] > maxDim) {
          maxDim = p.dim[d];
          t.splitRank = d;
        }
      }
      t.numSplit = 1;
      split_update(t, numMm, numMk, p);
      if (t.splitDim < 2) continue;
      const long long unsplit_bytes = (long long)t.volMmkUnsplit * B;
      const long long minSplit = (t.splitDim * unsplit_bytes + dev.lds_per_block - 1) / dev.lds_per_block;
      if (minSplit > t.splitDim || minSplit > 10000) continue;
      const int lo = (int)std::max<long long>(2, minSplit);
      const int hi = std::max(lo, std::min(t.splitDim / 2, lo + 60));
      // Keep: the first split count that launches, the one with the most
      // staged elements per CU, and the one with the best thread use.
      int best[3] = {0, 0, 0};
      long long bestv1 = -1, bestv2 = -1;
      for (int ns = lo; ns <= hi; ns++) {
        if ((long long)t.splitDim * (ns + 1) >= (1LL << 31)) break;
        Split u = t;
        u.numSplit = ns;
        split_update(u, numMm, numMk, p);
        const int chunk = (u.splitDim + ns - 1) / ns;
        const long long vs = (long long)chunk * u.volMmkUnsplit;
        if (vs * B > dev.lds_per_block) continue;
        if (!choose_threads((int)vs, B, dev, nt, nreg)) continue;
        const int act = std::max(1, std::min(dev.lds_per_cu / (int)(vs * B), dev.max_waves_per_cu / (nt / 64)));
        const long long used = (long long)(u.splitDim / ns) * u.volMmkUnsplit;
        const long long v1 = used * act;
        const long long v2 = (long long)nt * nreg > 0 ? used * 100 / ((long long)nt * nreg) : 0;
        if (!best[0]) best[0] = ns;
        if (v1 > bestv1) {
          bestv1 = v1;
          best[1] = ns;
        }
        if (v2 > bestv2) {
          bestv2 = v2;
          best[2] = ns;
        }
      }
      for (int k = 0; k < 3; k++) {
        if (!best[k]) continue;
        bool dup = false;
        for (int j = 0; j < k; j++) dup |= (best[j] == best[k]);
        if (dup) continue;
        Split u = t;
        u.numSplit = best[k];
        split_update(u, numMm, numMk, p);
        const int chunk = (u.splitDim + u.numSplit - 1) / u.numSplit;
        if (!choose_threads(chunk * u.volMmkUnsplit, B, dev, nt, nreg)) continue;
        push_packed(p, source, u, nt, nreg, dev, out);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// gfx950 cost model
// ---------------------------------------------------------------------------

struct ModelConst {
  double bw_frac = 0.80;        // sustained fraction of the HBM peak
  double partial_penalty = 1.0; // extra sector transfers per partially written sector
  double round_latency = 2.0e-6;// one load+store round of a resident block
  double launch = 4.0e-6;       // launch + completion overhead
};

const ModelConst kModel;

void model_tiled(Candidate& c, const DeviceInfo& dev) {
  const TiledArgs& a = c.tiled;
  const int B = c.elem;
  double R = 0, W = 0, P = 0;
  std::vector<long long> ws = sample_indices(a.items, 48);
  for (long long w : ws) {
    uint32_t mb = (uint32_t)(w / a.ntiles);
    int tile = (int)(w % a.ntiles);
    int x0 = (tile % a.ntx) * c.tile_x, y0 = (tile / a.ntx) * c.tile_y;
    int oi, oo;
    outer_offsets(a.outer, mb, oi, oo);
    int tx = std::min(c.tile_x, a.X - x0), ty = std::min(c.tile_y, a.Y - y0);
    for (int y = 0; y < ty; y++) R += sectors(oi + x0 + (long long)(y0 + y) * a.in_ys, tx, B);
    for (int x = 0; x < tx; x++) {
      long long s = oo + y0 + (long long)(x0 + x) * a.out_xs;
      W += sectors(s, ty, B);
      P += partial_sectors(s, ty, B);
    }
  }
  const double scale = ws.empty() ? 0 : (double)a.items / ws.size();
  c.model.read_sectors = R * scale;
  c.model.write_sectors = W * scale;
  c.model.partial_writes = P * scale;
  c.model.items = a.items;
  c.model.steps_per_block = (double)a.items / c.grid_x;
}

void model_copy(Candidate& c, const DeviceInfo& dev) {
  const CopyArgs& a = c.copy;
  const int Wb = c.word;
  const int ty_tile = (256 / c.copy_tx) * 4;
  double R = 0, W = 0, P = 0;
  std::vector<long long> ws = sample_indices(a.items, 48);
  for (long long w : ws) {
    uint32_t mb = (uint32_t)(w / a.ntiles);
    int tile = (int)(w % a.ntiles);
    int x0 = (tile % a.ntx) * c.copy_tx, y0 = (tile / a.ntx) * ty_tile;
    int oi, oo;
    outer_offsets(a.outer, mb, oi, oo);
    int tx = std::min(c.copy_tx, a.X - x0), ty = std::min(ty_tile, a.Y - y0);
    for (int y = 0; y < ty; y++) {
      R += sectors(oi + x0 + (long long)(y0 + y) * a.in_ys, tx, Wb);
      long long s = oo + x0 + (long long)(y0 + y) * a.out_ys;
      W += sectors(s, tx, Wb);
      P += partial_sectors(s, tx, Wb);
    }
  }
  const double scale = ws.empty() ? 0 : (double)a.items / ws.size();
  c.model.read_sectors = R * scale;
  c.model.write_sectors = W * scale;
  c.model.partial_writes = P * scale;
  c.model.items = a.items;
  c.model.steps_per_block = (double)a.items / c.grid_x;
}

// Contiguous runs (start offset, length) of one staged shape walked in LDS
// order (input side) or output order (output side).
void staged_runs(const StagedShape& s, bool output_side, std::vector<std::pair<long long, int>>& runs) {
  runs.clear();
  const int n = s.n;
  int dig[kMaxMmk] = {0};
  long long off = 0;
  long long run_start = 0;
  int run_len = 0;
  long long expect = 0;
  for (int p = 0; p < s.vol; p++) {
    if (run_len > 0 && off == expect) {
      run_len++;
    } else {
      if (run_len > 0) runs.push_back({run_start, run_len});
      run_start = off;
      run_len = 1;
    }
    expect = off + 1;
    for (int d = 0; d < n; d++) {
      const int dd = output_side ? (int)s.dout[d].d : (int)s.din[d].d;
      const int ct = output_side ? s.ct_out[d] : s.ct_in[d];
      dig[d]++;
      off += ct;
      if (dig[d] < dd) break;
      off -= (long long)dd * ct;
      dig[d] = 0;
    }
  }
  if (run_len > 0) runs.push_back({run_start, run_len});
}

void model_packed(Candidate& c, const DeviceInfo& dev) {
  const PackedArgs& a = c.packed;
  const int B = c.elem;
  double R = 0, W = 0, P = 0;
  std::vector<std::pair<long long, int>> rin[2], rout[2];
  const int nvar = (c.method == kPackedSplit) ? 2 : 1;
  for (int v = 0; v < nvar; v++) {
    staged_runs(a.shape[v], false, rin[v]);
    staged_runs(a.shape[v], true, rout[v]);
  }
  // Sample (split chunk, outer index) pairs.
  const long long nitems = (long long)a.num_split * a.vol_outer;
  std::vector<long long> ws = sample_indices(nitems, 8);
  for (long long w : ws) {
    int split = (int)(w % a.num_split);
    uint32_t o = (uint32_t)(w / a.num_split);
    const int q0 = a.split_dim / a.num_split;
    const int c0 = (int)((long long)split * a.split_dim / a.num_split);
    const int c1 = (int)((long long)(split + 1) * a.split_dim / a.num_split);
    const int plus = (nvar == 2) ? (c1 - c0) - q0 : 0;
    int oi, oo;
    outer_offsets(a.outer, o, oi, oo);
    const long long bi = (long long)c0 * a.split_in_stride + oi;
    const long long bo = (long long)c0 * a.split_out_stride + oo;
    for (const auto& r : rin[plus]) R += sectors(bi + r.first, r.second, B);
    for (const auto& r : rout[plus]) {
      W += sectors(bo + r.first, r.second, B);
      P += partial_sectors(bo + r.first, r.second, B);
    }
  }
  const double scale = ws.empty() ? 0 : (double)nitems / ws.size();
  c.model.read_sectors = R * scale;
  c.model.write_sectors = W * scale;
  c.model.partial_writes = P * scale;
  c.model.items = (double)nitems;
  c.model.steps_per_block = a.outer_per_block;
}

}  // namespace

void model_candidate(Candidate& c, const Problem& full, const DeviceInfo& dev) {
  ModelStats& m = c.model;
  m = ModelStats();
  if (c.method == kTrivial) {
    m.read_sectors = m.write_sectors = (double)(c.bytes / 2) / 64.0;
    m.t_hbm = (double)c.bytes / (dev.hbm_bytes_per_s * kModel.bw_frac);
    m.t_launch = kModel.launch;
    m.predicted_s = m.t_hbm + m.t_launch;
    return;
  }
  if (c.active_blocks_per_cu <= 0) c.active_blocks_per_cu = active_blocks(c, dev);
  switch (c.method) {
    case kTiled:
      model_tiled(c, dev);
      break;
    case kTiledCopy:
      model_copy(c, dev);
      break;
    default:
      model_packed(c, dev);
      break;
  }
  const double blocks = (double)c.grid_x * c.grid_y;
  m.concurrency = std::min(blocks, (double)dev.cu_count * std::max(1, c.active_blocks_per_cu));
  const double sector_bytes = 64.0 * (m.read_sectors + m.write_sectors + kModel.partial_penalty * m.partial_writes);
  // Bandwidth reachable with the resident blocks: each resident block keeps
  // about one round of its staged bytes in flight.
  const double bytes_per_round = (double)c.bytes / std::max(1.0, m.items);
  const double inflight = m.concurrency * bytes_per_round;
  const double bw_cap = inflight / kModel.round_latency;
  const double bw = std::min(dev.hbm_bytes_per_s * kModel.bw_frac, std::max(bw_cap, 1e9));
  m.t_hbm = sector_bytes / bw;
  m.t_latency = std::ceil(m.items / std::max(1.0, m.concurrency)) * kModel.round_latency * 0.5;
  m.t_launch = kModel.launch;
  m.predicted_s = std::max(m.t_hbm, m.t_latency) + m.t_launch;
  (void)full;
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
  add_packed(full, "full", dev, out);
  if (red.rank != full.rank) add_packed(red, "reduced", dev, out);
  for (auto& c : out) {
    if (c.active_blocks_per_cu <= 0) c.active_blocks_per_cu = active_blocks(c, dev);
    model_candidate(c, full, dev);
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

}  // namespace lrt
