// Persistent kernel of the batched engine, its shape table and host replay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "batch.h"

#include <cstdlib>
#include "kernel_ops.h"
#include "kernels.h"

namespace lrt {

namespace {

// Shape table: id, method, element bytes, four method parameters, group flag
//   Tiled:     tile_x, tile_y, vec_x, vec_y; group flag 1 for planes over
//              dim groups (their own family, so the one-dim kernels keep
//              their registers)
//   TiledCopy: word, copy_tx
//   Packed:    nreg
#define LRT_BATCH_TABLE(X)                                                  \
  X(0, kTiled, 4, 64, 64, 1, 1, 0) X(1, kTiled, 4, 64, 64, 4, 1, 0)         \
  X(2, kTiled, 4, 64, 64, 1, 4, 0) X(3, kTiled, 4, 64, 64, 4, 4, 0)         \
  X(4, kTiled, 8, 32, 32, 1, 1, 0) X(5, kTiled, 8, 32, 32, 2, 1, 0)         \
  X(6, kTiled, 8, 32, 32, 1, 2, 0) X(7, kTiled, 8, 32, 32, 2, 2, 0)         \
  X(8, kTiled, 16, 32, 32, 1, 1, 0)                                         \
  X(9, kTiledCopy, 4, 4, 16, 0, 0, 0) X(10, kTiledCopy, 4, 4, 64, 0, 0, 0)  \
  X(11, kTiledCopy, 8, 8, 16, 0, 0, 0) X(12, kTiledCopy, 8, 8, 64, 0, 0, 0) \
  X(13, kTiledCopy, 16, 16, 16, 0, 0, 0)                                    \
  X(14, kTiledCopy, 16, 16, 64, 0, 0, 0)                                    \
  X(15, kPacked, 4, 1, 0, 0, 0, 0) X(16, kPacked, 4, 2, 0, 0, 0, 0)         \
  X(17, kPacked, 4, 4, 0, 0, 0, 0) X(18, kPacked, 4, 8, 0, 0, 0, 0)         \
  X(19, kPacked, 8, 1, 0, 0, 0, 0) X(20, kPacked, 8, 2, 0, 0, 0, 0)         \
  X(21, kPacked, 8, 4, 0, 0, 0, 0) X(22, kPacked, 8, 8, 0, 0, 0, 0)         \
  X(23, kPacked, 16, 1, 0, 0, 0, 0) X(24, kPacked, 16, 2, 0, 0, 0, 0)       \
  X(25, kPacked, 16, 4, 0, 0, 0, 0)                                         \
  X(26, kTiled, 4, 64, 64, 1, 1, 1) X(27, kTiled, 4, 64, 64, 4, 1, 1)       \
  X(28, kTiled, 4, 64, 64, 1, 4, 1) X(29, kTiled, 4, 64, 64, 4, 4, 1)       \
  X(30, kTiled, 8, 32, 32, 1, 1, 1) X(31, kTiled, 8, 32, 32, 2, 1, 1)       \
  X(32, kTiled, 8, 32, 32, 1, 2, 1) X(33, kTiled, 8, 32, 32, 2, 2, 1)       \
  X(34, kTiled, 16, 32, 32, 1, 1, 1)

struct BatchShape {
  int id, method, b, p1, p2, p3, p4, g;
};
#define LRT_BATCH_ROW(ID, M, B, P1, P2, P3, P4, G) {ID, M, B, P1, P2, P3, P4, G},
const BatchShape kBatchShapes[] = {LRT_BATCH_TABLE(LRT_BATCH_ROW)};
#undef LRT_BATCH_ROW

// The problem owning unit u, searching forward from problem p (u only grows):
// the smallest q >= p with start[q + 1] > u. Each lane of the wave tests one
// problem, so one coalesced load and a ballot cover 64 problems (a scalar
// walk would take one dependent load per problem skipped, and a workgroup
// that strides by the grid skips many small problems per unit).
// start[nprob] = total > u ends the search inside the array.
__device__ __forceinline__ int batch_advance(const int* __restrict__ start, int nprob, int p, int u) {
  if (start[p + 1] > u) return p;
  const int lane = (int)(threadIdx.x & 63);
  for (int base = p + 1;; base += 64) {
    const int q = base + lane;
    const bool hit = q >= nprob || start[q + 1] > u;
    const unsigned long long m = __ballot(hit);
    if (m) return base + __ffsll((long long)m) - 1;
  }
}

__device__ __forceinline__ int batch_find(const int* __restrict__ start, int nprob, int u) {
  int lo = 0, hi = nprob;  // start[lo] <= u < start[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (start[mid] <= u)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

template <int B, int TX, int TY, int VX, int VY, bool G = false>
__device__ __forceinline__ void dev_tiled(const BatchDesc& D, int lu, unsigned char* lds_raw) {
  using Ops = TiledOps<B, TX, TY, VX, VY, G>;
  using E = typename Ops::E;
  E* lds = reinterpret_cast<E*>(lds_raw);
  for (int w = lu; w < D.items; w += D.groups) {
    const TileCtx c = Ops::item(D.a.tiled, (uint32_t)w);
    if constexpr (G) {
      Ops::prepare(D.a.tiled, c, threadIdx.x, lds);
      __syncthreads();
    }
    Ops::load(D.a.tiled, c, threadIdx.x, static_cast<const E*>(D.in), lds);
    __syncthreads();
    Ops::store(D.a.tiled, c, threadIdx.x, static_cast<E*>(D.out), lds);
    __syncthreads();
  }
}

template <int W, int TX>
__device__ __forceinline__ void dev_copy(const BatchDesc& D, int lu) {
  using Ops = CopyOps<W, TX, 4>;
  using V = typename Ops::V;
  for (int w = lu; w < D.items; w += D.groups) {
    const TileCtx c = Ops::item(D.a.copy, (uint32_t)w);
    Ops::run(D.a.copy, c, threadIdx.x, static_cast<const V*>(D.in), static_cast<V*>(D.out));
  }
}

template <int B, int NREG>
__device__ __forceinline__ void dev_packed(const BatchDesc& D, int lu, unsigned char* lds_raw) {
  using Ops = PackedOps<B, NREG>;
  using E = typename Ops::E;
  E* lds = reinterpret_cast<E*>(lds_raw);
  const PackedArgs& a = D.a.packed;
  const uint32_t og = fdiv((uint32_t)lu, D.split_div);
  const int split = lu - (int)(og * D.split_div.d);
  typename Ops::Pre p;
  Ops::prepare(a, split, threadIdx.x, 256, p);
  for (int o = (int)og; o < a.vol_outer; o += D.groups) {
    int off_in, off_out;
    outer_offsets(a.outer, (uint32_t)o, off_in, off_out);
    Ops::load(p, off_in, threadIdx.x, 256, static_cast<const E*>(D.in), lds);
    __syncthreads();
    Ops::store(p, off_out, threadIdx.x, 256, static_cast<E*>(D.out), lds);
    __syncthreads();
  }
}

// Runs unit lu of descriptor D if its shape belongs to family FAM (the
// switch of each family kernel holds only that family's shapes, which keeps
// its register use to what those shapes need).
template <int FAM>
__device__ __forceinline__ void dev_unit(const BatchDesc& D, int lu, unsigned char* lds) {
  if constexpr (FAM == kFamTiled) {
    switch (D.variant) {
      case 0: dev_tiled<4, 64, 64, 1, 1>(D, lu, lds); break;
      case 1: dev_tiled<4, 64, 64, 4, 1>(D, lu, lds); break;
      case 2: dev_tiled<4, 64, 64, 1, 4>(D, lu, lds); break;
      case 3: dev_tiled<4, 64, 64, 4, 4>(D, lu, lds); break;
      case 4: dev_tiled<8, 32, 32, 1, 1>(D, lu, lds); break;
      case 5: dev_tiled<8, 32, 32, 2, 1>(D, lu, lds); break;
      case 6: dev_tiled<8, 32, 32, 1, 2>(D, lu, lds); break;
      case 7: dev_tiled<8, 32, 32, 2, 2>(D, lu, lds); break;
      case 8: dev_tiled<16, 32, 32, 1, 1>(D, lu, lds); break;
      default: break;
    }
  } else if constexpr (FAM == kFamTiledG) {
    switch (D.variant) {
      case 26: dev_tiled<4, 64, 64, 1, 1, true>(D, lu, lds); break;
      case 27: dev_tiled<4, 64, 64, 4, 1, true>(D, lu, lds); break;
      case 28: dev_tiled<4, 64, 64, 1, 4, true>(D, lu, lds); break;
      case 29: dev_tiled<4, 64, 64, 4, 4, true>(D, lu, lds); break;
      case 30: dev_tiled<8, 32, 32, 1, 1, true>(D, lu, lds); break;
      case 31: dev_tiled<8, 32, 32, 2, 1, true>(D, lu, lds); break;
      case 32: dev_tiled<8, 32, 32, 1, 2, true>(D, lu, lds); break;
      case 33: dev_tiled<8, 32, 32, 2, 2, true>(D, lu, lds); break;
      case 34: dev_tiled<16, 32, 32, 1, 1, true>(D, lu, lds); break;
      default: break;
    }
  } else if constexpr (FAM == kFamCopy) {
    switch (D.variant) {
      case 9: dev_copy<4, 16>(D, lu); break;
      case 10: dev_copy<4, 64>(D, lu); break;
      case 11: dev_copy<8, 16>(D, lu); break;
      case 12: dev_copy<8, 64>(D, lu); break;
      case 13: dev_copy<16, 16>(D, lu); break;
      case 14: dev_copy<16, 64>(D, lu); break;
      default: break;
    }
  } else {
    switch (D.variant) {
      case 15: dev_packed<4, 1>(D, lu, lds); break;
      case 16: dev_packed<4, 2>(D, lu, lds); break;
      case 17: dev_packed<4, 4>(D, lu, lds); break;
      case 18: dev_packed<4, 8>(D, lu, lds); break;
      case 19: dev_packed<8, 1>(D, lu, lds); break;
      case 20: dev_packed<8, 2>(D, lu, lds); break;
      case 21: dev_packed<8, 4>(D, lu, lds); break;
      case 22: dev_packed<8, 8>(D, lu, lds); break;
      case 23: dev_packed<16, 1>(D, lu, lds); break;
      case 24: dev_packed<16, 2>(D, lu, lds); break;
      case 25: dev_packed<16, 4>(D, lu, lds); break;
      default: break;
    }
  }
}

// Persistent kernel of one family over `total` units of `nprob`
// descriptors: workgroup g takes units g, g + G, g + 2G, ...
template <int FAM>
__global__ __launch_bounds__(256, FAM == kFamPacked ? 6 : 1) void k_batch(const BatchDesc* __restrict__ descs, const int* __restrict__ start,
                                               int nprob, int total) {
  // TiledCopy moves data through registers only.
  __shared__ __attribute__((aligned(16))) unsigned char lds[FAM == kFamCopy ? 16 : kBatchLds];
  int u = blockIdx.x;
  if (u >= total) return;
  int p = batch_find(start, nprob, u);
  for (; u < total; u += gridDim.x) {
    p = batch_advance(start, nprob, p, u);
    dev_unit<FAM>(descs[p], u - start[p], lds);
  }
}

const void* batch_kernel(int fam) {
  switch (fam) {
    case kFamTiled: return reinterpret_cast<const void*>(&k_batch<kFamTiled>);
    case kFamTiledG: return reinterpret_cast<const void*>(&k_batch<kFamTiledG>);
    case kFamCopy: return reinterpret_cast<const void*>(&k_batch<kFamCopy>);
    default: return reinterpret_cast<const void*>(&k_batch<kFamPacked>);
  }
}

// ---------------------------------------------------------------------------
// Host replay: the same per-thread code, workgroup by workgroup.
// ---------------------------------------------------------------------------

template <int B, int TX, int TY, int VX, int VY, bool G = false>
void host_tiled(const BatchDesc& D, int lu, unsigned char* lds_raw) {
  using Ops = TiledOps<B, TX, TY, VX, VY, G>;
  using E = typename Ops::E;
  E* lds = reinterpret_cast<E*>(lds_raw);
  for (int w = lu; w < D.items; w += D.groups) {
    const TileCtx c = Ops::item(D.a.tiled, (uint32_t)w);
    for (int t = 0; t < 256; t++) Ops::prepare(D.a.tiled, c, t, lds);
    for (int t = 0; t < 256; t++) Ops::load(D.a.tiled, c, t, static_cast<const E*>(D.in), lds);
    for (int t = 0; t < 256; t++) Ops::store(D.a.tiled, c, t, static_cast<E*>(D.out), lds);
  }
}

template <int W, int TX>
void host_copy(const BatchDesc& D, int lu) {
  using Ops = CopyOps<W, TX, 4>;
  using V = typename Ops::V;
  for (int w = lu; w < D.items; w += D.groups) {
    const TileCtx c = Ops::item(D.a.copy, (uint32_t)w);
    for (int t = 0; t < 256; t++) Ops::run(D.a.copy, c, t, static_cast<const V*>(D.in), static_cast<V*>(D.out));
  }
}

template <int B, int NREG>
void host_packed(const BatchDesc& D, int lu, unsigned char* lds_raw) {
  using Ops = PackedOps<B, NREG>;
  using E = typename Ops::E;
  E* lds = reinterpret_cast<E*>(lds_raw);
  const PackedArgs& a = D.a.packed;
  const uint32_t og = fdiv((uint32_t)lu, D.split_div);
  const int split = lu - (int)(og * D.split_div.d);
  std::vector<typename Ops::Pre> pre(256);
  for (int t = 0; t < 256; t++) Ops::prepare(a, split, t, 256, pre[t]);
  for (int o = (int)og; o < a.vol_outer; o += D.groups) {
    int off_in, off_out;
    outer_offsets(a.outer, (uint32_t)o, off_in, off_out);
    for (int t = 0; t < 256; t++) Ops::load(pre[t], off_in, t, 256, static_cast<const E*>(D.in), lds);
    for (int t = 0; t < 256; t++) Ops::store(pre[t], off_out, t, 256, static_cast<E*>(D.out), lds);
  }
}

void host_unit(const BatchDesc& D, int lu, unsigned char* lds) {
  switch (D.variant) {
    case 0: host_tiled<4, 64, 64, 1, 1>(D, lu, lds); break;
    case 1: host_tiled<4, 64, 64, 4, 1>(D, lu, lds); break;
    case 2: host_tiled<4, 64, 64, 1, 4>(D, lu, lds); break;
    case 3: host_tiled<4, 64, 64, 4, 4>(D, lu, lds); break;
    case 4: host_tiled<8, 32, 32, 1, 1>(D, lu, lds); break;
    case 5: host_tiled<8, 32, 32, 2, 1>(D, lu, lds); break;
    case 6: host_tiled<8, 32, 32, 1, 2>(D, lu, lds); break;
    case 7: host_tiled<8, 32, 32, 2, 2>(D, lu, lds); break;
    case 8: host_tiled<16, 32, 32, 1, 1>(D, lu, lds); break;
    case 9: host_copy<4, 16>(D, lu); break;
    case 10: host_copy<4, 64>(D, lu); break;
    case 11: host_copy<8, 16>(D, lu); break;
    case 12: host_copy<8, 64>(D, lu); break;
    case 13: host_copy<16, 16>(D, lu); break;
    case 14: host_copy<16, 64>(D, lu); break;
    case 15: host_packed<4, 1>(D, lu, lds); break;
    case 16: host_packed<4, 2>(D, lu, lds); break;
    case 17: host_packed<4, 4>(D, lu, lds); break;
    case 18: host_packed<4, 8>(D, lu, lds); break;
    case 19: host_packed<8, 1>(D, lu, lds); break;
    case 20: host_packed<8, 2>(D, lu, lds); break;
    case 21: host_packed<8, 4>(D, lu, lds); break;
    case 22: host_packed<8, 8>(D, lu, lds); break;
    case 23: host_packed<16, 1>(D, lu, lds); break;
    case 24: host_packed<16, 2>(D, lu, lds); break;
    case 25: host_packed<16, 4>(D, lu, lds); break;
    case 26: host_tiled<4, 64, 64, 1, 1, true>(D, lu, lds); break;
    case 27: host_tiled<4, 64, 64, 4, 1, true>(D, lu, lds); break;
    case 28: host_tiled<4, 64, 64, 1, 4, true>(D, lu, lds); break;
    case 29: host_tiled<4, 64, 64, 4, 4, true>(D, lu, lds); break;
    case 30: host_tiled<8, 32, 32, 1, 1, true>(D, lu, lds); break;
    case 31: host_tiled<8, 32, 32, 2, 1, true>(D, lu, lds); break;
    case 32: host_tiled<8, 32, 32, 1, 2, true>(D, lu, lds); break;
    case 33: host_tiled<8, 32, 32, 2, 2, true>(D, lu, lds); break;
    case 34: host_tiled<16, 32, 32, 1, 1, true>(D, lu, lds); break;
    default: break;
  }
}

}  // namespace

bool batch_tiled_available(int elem, int tx, int ty, int vx, int vy, bool general) {
  for (const auto& s : kBatchShapes)
    if (s.method == kTiled && s.b == elem && s.p1 == tx && s.p2 == ty && s.p3 == vx && s.p4 == vy &&
        (s.g != 0) == general)
      return true;
  return false;
}

bool batch_copy_available(int word, int tx) {
  for (const auto& s : kBatchShapes)
    if (s.method == kTiledCopy && s.p1 == word && s.p2 == tx) return true;
  return false;
}

namespace {
// Bit nreg of the entry for elem / 4 (elements of 4, 8 and 16 bytes): the
// planner asks for every split it tries, so the table is indexed once.
struct PackedShapeBits {
  unsigned bits[5] = {0, 0, 0, 0, 0};
  PackedShapeBits() {
    for (const auto& s : kBatchShapes)
      if (s.method == kPacked && (s.b == 4 || s.b == 8 || s.b == 16) && s.p1 >= 0 && s.p1 < 32)
        bits[s.b / 4] |= 1u << s.p1;
  }
};
}  // namespace

bool batch_packed_available(int elem, int nreg) {
  static const PackedShapeBits t;
  if ((elem != 4 && elem != 8 && elem != 16) || nreg < 0 || nreg >= 32) return false;
  return (t.bits[elem / 4] >> nreg) & 1u;
}

int batch_variant(const Candidate& c) {
  for (const auto& s : kBatchShapes) {
    switch (c.method) {
      case kTiled:
        if (s.method == kTiled && s.b == c.elem && s.p1 == c.tile_x && s.p2 == c.tile_y && s.p3 == c.vec_x &&
            s.p4 == c.vec_y && (s.g != 0) == (c.xgroup > 1 || c.ygroup > 1))
          return s.id;
        break;
      case kTiledCopy:
        if (s.method == kTiledCopy && s.p1 == c.word && s.p2 == c.copy_tx) return s.id;
        break;
      case kPacked:
      case kPackedSplit:
        if (s.method == kPacked && s.b == c.elem && s.p1 == c.nreg && c.block == 256) return s.id;
        break;
      default:
        break;
    }
  }
  return -1;
}

bool batch_describe(const Candidate& c, BatchDesc& d) {
  d = BatchDesc();
  d.variant = batch_variant(c);
  if (d.variant < 0) return false;
  if (c.lds_bytes > kBatchLds) return false;
  d.split_div = make_fastdiv(1);
  switch (c.method) {
    case kTiled:
      d.a.tiled = c.tiled;
      d.items = c.tiled.items;
      d.item_bytes = 2 * c.tile_x * c.tile_y * c.elem;
      break;
    case kTiledCopy:
      d.a.copy = c.copy;
      d.items = c.copy.items;
      d.item_bytes = 2 * 1024 * c.word;  // copy_tx words x 1024/copy_tx rows
      break;
    default:
      d.a.packed = c.packed;
      d.items = c.packed.vol_outer;
      d.split_div = make_fastdiv((uint32_t)c.packed.num_split);
      d.item_bytes = std::max(1, 2 * std::max(1, c.packed.shape[0].vol) * c.elem / std::max(1, c.packed.num_split));
      break;
  }
  if (d.items <= 0) return false;
  batch_regroup(d, kBatchUnitBytes);
  return d.units > 0;
}

void batch_regroup(BatchDesc& d, long long unit_bytes) {
  const long long r = std::max<long long>(1, std::min<long long>(d.items, unit_bytes / std::max(1, d.item_bytes)));
  d.rounds = (int)r;
  d.groups = (int)((d.items + r - 1) / r);
  d.units = d.groups * (int)d.split_div.d;
}

long long batch_unit_bytes(long long total_bytes, int grid) {
  // LIBRETT_BATCH_UNIT_KB fixes the unit size; LIBRETT_BATCH_UNITS_PER_WG
  // sets the units each workgroup should get (default 8).
  static const long long fixed = [] {
    const char* e = std::getenv("LIBRETT_BATCH_UNIT_KB");
    return e ? std::max(1LL, std::atoll(e)) << 10 : 0LL;
  }();
  static const long long per_wg = [] {
    const char* e = std::getenv("LIBRETT_BATCH_UNITS_PER_WG");
    return e ? std::max(1LL, std::atoll(e)) : 32LL;
  }();
  if (fixed) return fixed;
  const long long per = total_bytes / (per_wg * std::max(1, grid));
  return std::max<long long>(16 << 10, std::min<long long>(kBatchUnitBytes, per));
}

int batch_family(int variant) {
  return variant <= 8 ? kFamTiled : variant <= 14 ? kFamCopy : variant <= 25 ? kFamPacked : kFamTiledG;
}

hipError_t launch_batch(int fam, const BatchDesc* d_descs, const int* d_start, int nprob, int total, int grid,
                        hipStream_t stream) {
  if (total <= 0 || nprob <= 0) return hipSuccess;
  grid = std::max(1, std::min(grid, total));
  void* args[4] = {&d_descs, &d_start, &nprob, &total};
  return hipLaunchKernel(batch_kernel(fam), dim3(grid), dim3(256), args, 0, stream);
}

int batch_occupancy(int fam) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, batch_kernel(fam), 256, 0) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

void batch_layout(std::vector<BatchDesc>& descs, BatchLayout& L, std::vector<int>* order) {
  const int nd = (int)descs.size();
  std::vector<int> idx(nd);
  for (int k = 0; k < nd; k++) idx[k] = k;
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) {
    return batch_family(descs[a].variant) < batch_family(descs[b].variant);
  });
  std::vector<BatchDesc> sorted(nd);
  for (int k = 0; k < nd; k++) sorted[k] = descs[idx[k]];
  descs.swap(sorted);
  if (order) *order = idx;
  L.start.assign(nd + kBatchFamilies, 0);
  int k = 0, off = 0;
  for (int f = 0; f < kBatchFamilies; f++) {
    L.first[f] = k;
    L.start_off[f] = off;
    long long tot = 0;
    int n = 0;
    while (k < nd && batch_family(descs[k].variant) == f) {
      L.start[off + n] = (int)std::min<long long>(tot, 0x7fffffff);
      tot += descs[k].units;
      n++;
      k++;
    }
    L.start[off + n] = (int)std::min<long long>(tot, 0x7fffffff);
    L.n[f] = n;
    L.total[f] = tot;
    off += n + 1;
  }
}

hipError_t launch_batch_all(const BatchDesc* d_descs, const int* d_start, const BatchLayout& L, const int* grid,
                            hipStream_t stream) {
  for (int f = 0; f < kBatchFamilies; f++) {
    if (L.n[f] == 0) continue;
    const hipError_t e =
        launch_batch(f, d_descs + L.first[f], d_start + L.start_off[f], L.n[f], (int)L.total[f], grid[f], stream);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

bool replay_batch_all(const BatchDesc* descs, const BatchLayout& L, int grid) {
  for (int f = 0; f < kBatchFamilies; f++) {
    if (L.n[f] == 0) continue;
    if (!replay_batch(descs + L.first[f], L.start.data() + L.start_off[f], L.n[f], (int)L.total[f], grid))
      return false;
  }
  return true;
}

bool replay_batch(const BatchDesc* descs, const int* start, int nprob, int total, int grid) {
  // The workgroup's LDS, followed by a guard band that no unit may touch.
  const size_t guard = 64 << 10;
  std::vector<unsigned char> lds(kBatchLds + guard, 0xC3);
  for (int g = 0; g < grid; g++) {
    int u = g;
    if (u >= total) continue;
    int p = 0;
    while (p < nprob && start[p + 1] <= u) p++;
    for (; u < total; u += grid) {
      while (start[p + 1] <= u) p++;
      host_unit(descs[p], u - start[p], lds.data());
    }
  }
  for (size_t i = kBatchLds; i < lds.size(); i++)
    if (lds[i] != 0xC3) return false;
  return true;
}

}  // namespace lrt
