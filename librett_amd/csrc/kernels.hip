// gfx950 transpose kernels, their launch table, runtime occupancy queries and
// the host replay used by the CPU tests.
//
// Reference counterpart of the dispatch: librettKernel, src/kernel.cpp:1008-1215
// (method x element type x register-count switch, src/calls.h:1-92).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "kernel_ops.h"
#include "kernels.h"

// k_packed with 12 or more elements per thread fetches the next round's
// input into registers while it stores the current round from LDS. Such
// workgroups are large and few per CU, and without the overlap each round
// waits out its full load latency: on MI355X the pipelined form ran these
// plans up to 1.8x faster, while for 2-8 elements per thread it was within
// a few percent either way (slower on the bench's Packed cases), so those
// fetch each round after the previous one is stored. For 8-byte elements
// with 16 per thread the extra registers no longer fit the 128 VGPRs of a
// 1024-thread workgroup and spill, so that instantiation is not pipelined.
// LRT_PACKED_PIPE=0 turns the pipeline off everywhere.
#ifndef LRT_PACKED_PIPE
#define LRT_PACKED_PIPE 1
#endif
template <int B, int NREG>
constexpr bool packed_pipelined() {
  return LRT_PACKED_PIPE && NREG >= 12 && !(B == 8 && NREG == 16);
}

namespace lrt {

// ---------------------------------------------------------------------------
// Device kernels
// ---------------------------------------------------------------------------

// k_tiled walks its tiles with a stride of the grid, one LDS round per
// tile. (Fetching the next tile into registers during the current tile's
// store phase, as k_packed does for its large rounds, measured slower on
// MI355X: 2-3% on the 8192 x 8192 planes and far slower on tiles whose
// registers then cut occupancy, so Tiled rounds are not pipelined.)
template <int B, int TX, int TY, int VX, int VY, bool G>
__global__ __launch_bounds__(256) void k_tiled(TiledArgs a, const void* __restrict__ in,
                                               void* __restrict__ out) {
  using Ops = TiledOps<B, TX, TY, VX, VY, G>;
  using E = typename Ops::E;
  __shared__ E lds[Ops::LDS_ELEMS];
  const E* src = static_cast<const E*>(in);
  E* dst = static_cast<E*>(out);
  const uint32_t w0 = a.xcds > 1 ? (uint32_t)xcd_logical((int)blockIdx.x, (int)gridDim.x, a.xcds) : blockIdx.x;
  for (uint32_t w = w0; w < (uint32_t)a.items; w += gridDim.x) {
    const TileCtx c = Ops::item(a, w);
    if constexpr (G) {
      Ops::prepare(a, c, threadIdx.x, lds);
      __syncthreads();
    }
    Ops::load(a, c, threadIdx.x, src, lds);
    __syncthreads();
    Ops::store(a, c, threadIdx.x, dst, lds);
    __syncthreads();
  }
}

template <int W, int TX, int NP>
__global__ __launch_bounds__(256) void k_copy(CopyArgs a, const void* __restrict__ in,
                                              void* __restrict__ out) {
  using Ops = CopyOps<W, TX, NP>;
  using V = typename Ops::V;
  const V* src = static_cast<const V*>(in);
  V* dst = static_cast<V*>(out);
  const uint32_t w0 = a.xcds > 1 ? (uint32_t)xcd_logical((int)blockIdx.x, (int)gridDim.x, a.xcds) : blockIdx.x;
  for (uint32_t w = w0; w < (uint32_t)a.items; w += gridDim.x) {
    const TileCtx c = Ops::item(a, w);
    Ops::run(a, c, threadIdx.x, src, dst);
  }
}

template <int B, int NREG>
__global__ __launch_bounds__(1024) void k_packed(PackedArgs a, const void* __restrict__ in,
                                                 void* __restrict__ out) {
  using Ops = PackedOps<B, NREG>;
  using E = typename Ops::E;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  E* lds = reinterpret_cast<E*>(smem);
  const E* src = static_cast<const E*>(in);
  E* dst = static_cast<E*>(out);
  const int nt = blockDim.x;
  const int t = threadIdx.x;
  typename Ops::Pre p;
  int split, o0, o1, step;
  if (a.order == 0) {
    split = blockIdx.x;
    o0 = blockIdx.y * a.outer_per_block;
    o1 = min(a.vol_outer, o0 + a.outer_per_block);
    step = 1;
  } else {
    const int G = gridDim.x * gridDim.y;
    const int L = xcd_logical(blockIdx.x + gridDim.x * blockIdx.y, G, a.xcds);
    split = L / gridDim.y;
    o0 = L - split * gridDim.y;
    o1 = a.vol_outer;
    step = gridDim.y;
  }
  Ops::prepare(a, split, t, nt, p);
  if constexpr (packed_pipelined<B, NREG>()) {
    // Round o is stashed from registers into LDS; while it is stored from
    // LDS, round o + step is fetched into the same registers.
    E r[NREG];
    if (o0 < o1) {
      int off_in, off_out;
      outer_offsets(a.outer, (uint32_t)o0, off_in, off_out);
      Ops::fetch(p, off_in, t, nt, src, r);
    }
    for (int o = o0; o < o1; o += step) {
      int off_in, off_out;
      outer_offsets(a.outer, (uint32_t)o, off_in, off_out);
      Ops::stash(p, t, nt, r, lds);
      __syncthreads();
      if (o + step < o1) {
        int next_in, next_out;
        outer_offsets(a.outer, (uint32_t)(o + step), next_in, next_out);
        Ops::fetch(p, next_in, t, nt, src, r);
      }
      Ops::store(p, off_out, t, nt, dst, lds);
      __syncthreads();
    }
  } else {
    for (int o = o0; o < o1; o += step) {
      int off_in, off_out;
      outer_offsets(a.outer, (uint32_t)o, off_in, off_out);
      Ops::load(p, off_in, t, nt, src, lds);
      __syncthreads();
      Ops::store(p, off_out, t, nt, dst, lds);
      __syncthreads();
    }
  }
}

// Plain device copy used as a bandwidth ceiling (the reference's bench copy
// kernels, src/GpuMemcpy.cpp:42-209, serve the same purpose). Each thread
// moves K 16-byte words per round, all loads issued before the stores.
template <int K, bool NT>
__global__ __launch_bounds__(256) void k_devcopy(const u32x4* __restrict__ in, u32x4* __restrict__ out,
                                                 long long n16) {
  const long long stride = (long long)gridDim.x * 256 * K;
  for (long long b = (long long)blockIdx.x * 256 * K + threadIdx.x; b < n16; b += stride) {
    u32x4 r[K];
#pragma unroll
    for (int k = 0; k < K; k++)
      if (b + k * 256 < n16) r[k] = in[b + k * 256];
#pragma unroll
    for (int k = 0; k < K; k++)
      if (b + k * 256 < n16) {
        if (NT)
          __builtin_nontemporal_store(r[k], out + b + k * 256);
        else
          out[b + k * 256] = r[k];
      }
  }
}

hipError_t launch_device_copy(void* dst, const void* src, size_t bytes, hipStream_t stream, int k, int nt,
                              int grid) {
  if (bytes % 16 != 0 || ((uintptr_t)dst | (uintptr_t)src) % 16 != 0) return hipErrorInvalidValue;
  const long long n16 = (long long)(bytes / 16);
  if (n16 == 0) return hipSuccess;
  const long long per_block = 256LL * k;
  long long need = (n16 + per_block - 1) / per_block;
  if (grid <= 0 || grid > need) grid = (int)std::min<long long>(need, 1 << 30);
  const void* kern = nullptr;
#define LRT_DC(K_, NT_) \
  if (k == K_ && nt == NT_) kern = reinterpret_cast<const void*>(&k_devcopy<K_, NT_ != 0>);
  LRT_DC(1, 0) LRT_DC(2, 0) LRT_DC(4, 0) LRT_DC(8, 0) LRT_DC(1, 1) LRT_DC(2, 1) LRT_DC(4, 1) LRT_DC(8, 1)
#undef LRT_DC
  if (!kern) return hipErrorInvalidValue;
  void* args[3] = {&src, &dst, const_cast<long long*>(&n16)};
  return hipLaunchKernel(kern, dim3(grid), dim3(256), args, 0, stream);
}

void kernels_warm_up() {
  hipFuncAttributes attr;
  if (hipFuncGetAttributes(&attr, reinterpret_cast<const void*>(&k_devcopy<1, false>)) != hipSuccess)
    (void)hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host replays: the same per-thread code, barrier order kept by running every
// thread's load() before any thread's store().
// ---------------------------------------------------------------------------

thread_local ReplayObserver* g_replay_observer = nullptr;

static inline void replay_at(long long item, int t) {
  if (g_replay_observer) g_replay_observer->at(item, t);
}

template <int B, int TX, int TY, int VX, int VY, bool G>
static void replay_tiled(const Candidate& c, const void* in, void* out) {
  using Ops = TiledOps<B, TX, TY, VX, VY, G>;
  using E = typename Ops::E;
  std::vector<E> lds(Ops::LDS_ELEMS);
  const TiledArgs& a = c.tiled;
  const E* src = static_cast<const E*>(in);
  E* dst = static_cast<E*>(out);
  for (uint32_t w = 0; w < (uint32_t)a.items; w++) {
    const TileCtx ctx = Ops::item(a, w);
    for (int t = 0; t < Ops::NT; t++) Ops::prepare(a, ctx, t, lds.data());
    for (int t = 0; t < Ops::NT; t++) {
      replay_at(w, t);
      Ops::load(a, ctx, t, src, lds.data());
    }
    for (int t = 0; t < Ops::NT; t++) {
      replay_at(w, t);
      Ops::store(a, ctx, t, dst, lds.data());
    }
  }
}

template <int W, int TX, int NP>
static void replay_copy(const Candidate& c, const void* in, void* out) {
  using Ops = CopyOps<W, TX, NP>;
  using V = typename Ops::V;
  const CopyArgs& a = c.copy;
  for (uint32_t w = 0; w < (uint32_t)a.items; w++) {
    const TileCtx ctx = Ops::item(a, w);
    for (int t = 0; t < Ops::NT; t++) {
      replay_at(w, t);
      Ops::run(a, ctx, t, static_cast<const V*>(in), static_cast<V*>(out));
    }
  }
}

template <int B, int NREG>
static void replay_packed(const Candidate& c, const void* in, void* out) {
  using Ops = PackedOps<B, NREG>;
  using E = typename Ops::E;
  const PackedArgs& a = c.packed;
  const int nt = c.block;
  std::vector<E> lds(c.lds_bytes / B + 1);
  std::vector<typename Ops::Pre> pre(nt);
  const int G = c.grid_x * c.grid_y;
  for (int b = 0; b < G; b++) {
    const int bx = b % c.grid_x, by = b / c.grid_x;
    int split, o0, o1, step;
    if (a.order == 0) {
      split = bx;
      o0 = by * a.outer_per_block;
      o1 = std::min(a.vol_outer, o0 + a.outer_per_block);
      step = 1;
    } else {
      const int L = xcd_logical(b, G, a.xcds);
      split = L / c.grid_y;
      o0 = L - split * c.grid_y;
      o1 = a.vol_outer;
      step = c.grid_y;
    }
    for (int t = 0; t < nt; t++) Ops::prepare(a, split, t, nt, pre[t]);
    const E* src = static_cast<const E*>(in);
    E* dst = static_cast<E*>(out);
    if constexpr (packed_pipelined<B, NREG>()) {
      // The device's schedule: fetch round o + step between the barrier
      // after round o's stash and round o's store.
      std::vector<E> regs((size_t)nt * NREG);
      if (o0 < o1) {
        int off_in, off_out;
        outer_offsets(a.outer, (uint32_t)o0, off_in, off_out);
        for (int t = 0; t < nt; t++) {
          replay_at((long long)split * a.vol_outer + o0, t);
          Ops::fetch(pre[t], off_in, t, nt, src, &regs[(size_t)t * NREG]);
        }
      }
      for (int o = o0; o < o1; o += step) {
        int off_in, off_out;
        outer_offsets(a.outer, (uint32_t)o, off_in, off_out);
        const long long item = (long long)split * a.vol_outer + o;
        for (int t = 0; t < nt; t++) Ops::stash(pre[t], t, nt, &regs[(size_t)t * NREG], lds.data());
        if (o + step < o1) {
          int next_in, next_out;
          outer_offsets(a.outer, (uint32_t)(o + step), next_in, next_out);
          for (int t = 0; t < nt; t++) {
            replay_at(item + step, t);
            Ops::fetch(pre[t], next_in, t, nt, src, &regs[(size_t)t * NREG]);
          }
        }
        for (int t = 0; t < nt; t++) {
          replay_at(item, t);
          Ops::store(pre[t], off_out, t, nt, dst, lds.data());
        }
      }
    } else {
      for (int o = o0; o < o1; o += step) {
        int off_in, off_out;
        outer_offsets(a.outer, (uint32_t)o, off_in, off_out);
        const long long item = (long long)split * a.vol_outer + o;
        for (int t = 0; t < nt; t++) {
          replay_at(item, t);
          Ops::load(pre[t], off_in, t, nt, src, lds.data());
        }
        for (int t = 0; t < nt; t++) {
          replay_at(item, t);
          Ops::store(pre[t], off_out, t, nt, dst, lds.data());
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Instantiation tables
// ---------------------------------------------------------------------------

struct TiledEntry {
  int b, tx, ty, vx, vy;
  bool general;  // handles x / y groups of several dims
  const void* kernel;
  void (*replay)(const Candidate&, const void*, void*);
};
struct CopyEntry {
  int w, tx, np;
  const void* kernel;
  void (*replay)(const Candidate&, const void*, void*);
};
struct PackedEntry {
  int b, nreg;
  const void* kernel;
  void (*replay)(const Candidate&, const void*, void*);
};

#define LRT_TILED_G(B, TX, TY, VX, VY, G)                                                    \
  {B, TX, TY, VX, VY, G, reinterpret_cast<const void*>(&k_tiled<B, TX, TY, VX, VY, G>), \
   &replay_tiled<B, TX, TY, VX, VY, G>},
#define LRT_TILED(B, TX, TY, VX, VY) LRT_TILED_G(B, TX, TY, VX, VY, false)
#define LRT_TILED_V(B, TX, TY, V) \
  LRT_TILED(B, TX, TY, 1, 1) LRT_TILED(B, TX, TY, V, 1) LRT_TILED(B, TX, TY, 1, V) LRT_TILED(B, TX, TY, V, V)
#define LRT_TILED_GV(B, TX, TY, V)                                                           \
  LRT_TILED_G(B, TX, TY, 1, 1, true) LRT_TILED_G(B, TX, TY, V, 1, true) LRT_TILED_G(B, TX, TY, 1, V, true) \
  LRT_TILED_G(B, TX, TY, V, V, true)

#define LRT_TILED_2(TX, TY) \
  LRT_TILED(4, TX, TY, 2, 1) LRT_TILED(4, TX, TY, 1, 2) LRT_TILED(4, TX, TY, 2, 2) LRT_TILED(4, TX, TY, 4, 2) \
  LRT_TILED(4, TX, TY, 2, 4)

static const TiledEntry kTiledTable[] = {
    LRT_TILED_V(4, 64, 64, 4) LRT_TILED_V(4, 32, 32, 4) LRT_TILED_V(4, 64, 32, 4) LRT_TILED_V(4, 32, 64, 4)
    LRT_TILED_V(8, 32, 32, 2) LRT_TILED_V(8, 64, 64, 2) LRT_TILED_V(8, 64, 32, 2) LRT_TILED_V(8, 32, 64, 2)
    // Tiles with 512-byte row segments (66 KiB of LDS for the largest).
    LRT_TILED_V(4, 128, 128, 4) LRT_TILED_V(4, 128, 64, 4) LRT_TILED_V(4, 64, 128, 4)
    LRT_TILED_V(8, 128, 64, 2) LRT_TILED_V(8, 64, 128, 2)
    LRT_TILED(16, 16, 16, 1, 1) LRT_TILED(16, 32, 32, 1, 1) LRT_TILED(16, 32, 16, 1, 1) LRT_TILED(16, 16, 32, 1, 1)
    // f32 with 8-byte lanes on a side whose extent is even but not a
    // multiple of 4 (half the wave instructions of 4-byte lanes).
    LRT_TILED_2(64, 64) LRT_TILED_2(32, 32) LRT_TILED_2(64, 32) LRT_TILED_2(32, 64) LRT_TILED_2(128, 64)
    LRT_TILED_2(64, 128)
    // Planes whose x or y is a group of several dims (separate
    // instantiations, so that the one-dim kernels above keep their
    // registers and occupancy).
    LRT_TILED_GV(4, 64, 64, 4) LRT_TILED_GV(4, 32, 32, 4) LRT_TILED_GV(4, 128, 64, 4) LRT_TILED_GV(4, 64, 128, 4)
    LRT_TILED_GV(8, 64, 64, 2) LRT_TILED_GV(8, 32, 32, 2) LRT_TILED_GV(8, 128, 64, 2) LRT_TILED_GV(8, 64, 128, 2)
    LRT_TILED_G(16, 16, 16, 1, 1, true) LRT_TILED_G(16, 32, 32, 1, 1, true)
};

#define LRT_COPY(W, TX) \
  {W, TX, 4, reinterpret_cast<const void*>(&k_copy<W, TX, 4>), &replay_copy<W, TX, 4>},
#define LRT_COPY_W(W) LRT_COPY(W, 4) LRT_COPY(W, 8) LRT_COPY(W, 16) LRT_COPY(W, 32) LRT_COPY(W, 64)

static const CopyEntry kCopyTable[] = {LRT_COPY_W(4) LRT_COPY_W(8) LRT_COPY_W(16)};

#define LRT_PACKED(B, N) {B, N, reinterpret_cast<const void*>(&k_packed<B, N>), &replay_packed<B, N>},
#define LRT_PACKED_8(B) \
  LRT_PACKED(B, 1) LRT_PACKED(B, 2) LRT_PACKED(B, 3) LRT_PACKED(B, 4) LRT_PACKED(B, 5) LRT_PACKED(B, 6) \
  LRT_PACKED(B, 7) LRT_PACKED(B, 8)

static const PackedEntry kPackedTable[] = {
    LRT_PACKED_8(4) LRT_PACKED(4, 12) LRT_PACKED(4, 16)
    LRT_PACKED_8(8) LRT_PACKED(8, 12) LRT_PACKED(8, 16)
    LRT_PACKED_8(16)
};

static const TiledEntry* find_tiled(const Candidate& c) {
  const bool general = c.xgroup > 1 || c.ygroup > 1;
  for (const auto& e : kTiledTable)
    if (e.b == c.elem && e.tx == c.tile_x && e.ty == c.tile_y && e.vx == c.vec_x && e.vy == c.vec_y &&
        e.general == general)
      return &e;
  return nullptr;
}
static const CopyEntry* find_copy(const Candidate& c) {
  for (const auto& e : kCopyTable)
    if (e.w == c.word && e.tx == c.copy_tx) return &e;
  return nullptr;
}
static const PackedEntry* find_packed(const Candidate& c) {
  for (const auto& e : kPackedTable)
    if (e.b == c.elem && e.nreg == c.nreg) return &e;
  return nullptr;
}

bool tiled_shape_available(int elem, int tx, int ty, int vx, int vy, bool general) {
  for (const auto& e : kTiledTable)
    if (e.b == elem && e.tx == tx && e.ty == ty && e.vx == vx && e.vy == vy && e.general == general) return true;
  return false;
}
bool copy_shape_available(int word, int tx) {
  for (const auto& e : kCopyTable)
    if (e.w == word && e.tx == tx) return true;
  return false;
}
int packed_max_nreg(int elem) {
  int m = 0;
  for (const auto& e : kPackedTable)
    if (e.b == elem && e.nreg > m) m = e.nreg;
  return m;
}
bool packed_nreg_available(int elem, int nreg) {
  for (const auto& e : kPackedTable)
    if (e.b == elem && e.nreg == nreg) return true;
  return false;
}

static const void* kernel_of(const Candidate& c) {
  switch (c.method) {
    case kTiled: {
      const TiledEntry* e = find_tiled(c);
      return e ? e->kernel : nullptr;
    }
    case kTiledCopy: {
      const CopyEntry* e = find_copy(c);
      return e ? e->kernel : nullptr;
    }
    case kPacked:
    case kPackedSplit: {
      const PackedEntry* e = find_packed(c);
      return e ? e->kernel : nullptr;
    }
    default:
      return nullptr;
  }
}

hipError_t launch_candidate(const Candidate& c, const void* in, void* out, hipStream_t stream) {
  if (c.method == kTrivial) {
    return hipMemcpyAsync(out, in, (size_t)(c.bytes / 2), hipMemcpyDeviceToDevice, stream);
  }
  const void* k = kernel_of(c);
  if (!k) return hipErrorInvalidDeviceFunction;
  dim3 grid(c.grid_x, c.grid_y, 1);
  dim3 block(c.block, 1, 1);
  void* args[3];
  switch (c.method) {
    case kTiled:
      args[0] = const_cast<TiledArgs*>(&c.tiled);
      break;
    case kTiledCopy:
      args[0] = const_cast<CopyArgs*>(&c.copy);
      break;
    default:
      args[0] = const_cast<PackedArgs*>(&c.packed);
      break;
  }
  args[1] = &in;
  args[2] = &out;
  return hipLaunchKernel(k, grid, block, args, (size_t)((c.method == kPacked || c.method == kPackedSplit) ? c.lds_bytes : 0),
                         stream);
}

bool replay_candidate(const Candidate& c, const void* in, void* out) {
  switch (c.method) {
    case kTrivial:
      std::memcpy(out, in, (size_t)(c.bytes / 2));
      return true;
    case kTiled: {
      const TiledEntry* e = find_tiled(c);
      if (!e) return false;
      e->replay(c, in, out);
      return true;
    }
    case kTiledCopy: {
      const CopyEntry* e = find_copy(c);
      if (!e) return false;
      e->replay(c, in, out);
      return true;
    }
    case kPacked:
    case kPackedSplit: {
      const PackedEntry* e = find_packed(c);
      if (!e) return false;
      e->replay(c, in, out);
      return true;
    }
    default:
      return false;
  }
}

// The dynamic-LDS limit of a kernel is a process-wide attribute of that
// kernel, shared by every plan that uses it. It is only ever raised, and
// raised straight to the device's per-workgroup maximum when the runtime
// accepts that, so no plan can lower the limit another plan's launches rely
// on, and planning threads cannot race it downward. Launches above the
// default limit need it.
static std::mutex g_lds_mu;
static std::map<std::pair<int, const void*>, int> g_lds_limit;

static bool ensure_lds_limit(const void* k, int bytes) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  std::lock_guard<std::mutex> lock(g_lds_mu);
  int& cur = g_lds_limit[{dev, k}];
  if (cur >= bytes) return true;
  int maxb = 0;
  if (hipDeviceGetAttribute(&maxb, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
    (void)hipGetLastError();
    maxb = 0;
  }
  if (maxb >= bytes && hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, maxb) == hipSuccess) {
    cur = maxb;
    return true;
  }
  (void)hipGetLastError();
  if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) == hipSuccess) {
    cur = bytes;
    return true;
  }
  (void)hipGetLastError();
  return false;
}

// Occupancy answers per (device, kernel, block, dynamic LDS): planning asks
// for the same few combinations over and over (the reference keeps the same
// kind of cache, nabCache, src/kernel.cpp:671-772; its key confuses 8- and
// 16-byte kernels, this one names the kernel itself).
struct OccKey {
  int dev;
  const void* k;
  int block, lds;
  bool operator<(const OccKey& o) const {
    if (dev != o.dev) return dev < o.dev;
    if (k != o.k) return k < o.k;
    if (block != o.block) return block < o.block;
    return lds < o.lds;
  }
};
static std::mutex g_occ_mu;
static std::map<OccKey, int> g_occ;

int query_occupancy(const Candidate& c) {
  const void* k = kernel_of(c);
  if (!k) return 0;
  const bool dyn = c.method == kPacked || c.method == kPackedSplit;
  const int lds = dyn ? c.lds_bytes : 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  const OccKey key{dev, k, c.block, lds};
  {
    std::lock_guard<std::mutex> lock(g_occ_mu);
    auto it = g_occ.find(key);
    if (it != g_occ.end()) return it->second;
  }
  if (dyn && !ensure_lds_limit(k, lds)) return 0;
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, c.block, (size_t)lds) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  std::lock_guard<std::mutex> lock(g_occ_mu);
  g_occ[key] = n;
  return n;
}

int prepare_kernel(const Candidate& c) {
  if (c.method != kPacked && c.method != kPackedSplit) return 0;
  const void* k = kernel_of(c);
  if (!k) return -1;
  return ensure_lds_limit(k, c.lds_bytes) ? 0 : -1;
}

}  // namespace lrt
