// Per-thread work of the transpose kernels, written once for the device and
// for the host replay used by the CPU tests.
//
// Each kernel is a sequence of "rounds": every thread of a workgroup runs
// load(), the workgroup synchronises, every thread runs store(). The device
// kernels in kernels.hip wrap these with __syncthreads(); the host replay runs
// load() for all threads of a block, then store() for all threads, which is
// exactly the ordering the barrier guarantees.
#pragma once

#include "common.h"

namespace lrt {

typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Opaque B-byte words: element moves never interpret the bits.
template <int BYTES> struct Word;
template <> struct Word<4> { using type = uint32_t; };
template <> struct Word<8> { using type = u32x2; };
template <> struct Word<16> { using type = u32x4; };

template <typename E, typename V>
__host__ __device__ inline E vget(const V& v, int k) {
  E e;
  __builtin_memcpy(&e, reinterpret_cast<const char*>(&v) + k * sizeof(E), sizeof(E));
  return e;
}

template <typename E, typename V>
__host__ __device__ inline void vset(V& v, int k, const E& e) {
  __builtin_memcpy(reinterpret_cast<char*>(&v) + k * sizeof(E), &e, sizeof(E));
}

// Global stores and loads of the transpose kernels. Every output byte is
// written once and never read back by the kernel, and every input byte is
// read once, so streaming (non-temporal) accesses keep the caches from
// holding lines nobody will reuse. On MI355X, non-temporal stores lift a
// plain 256 MiB copy from about 5.5 to about 7.1 TB/s. LRT_NT_STORES and
// LRT_NT_LOADS select them at build time.
#ifndef LRT_NT_STORES
#define LRT_NT_STORES 1
#endif
#ifndef LRT_NT_LOADS
#define LRT_NT_LOADS 0
#endif
// Host replay only: an observer that is told which work item and thread the
// replay is running and sees every global load and store it makes. The
// tests of the access counters compare these accesses with the counters'
// own enumeration. Null (the normal case) costs one test per access.
struct ReplayObserver {
  virtual void at(long long item, int thread) = 0;
  virtual void access(const void* p, int bytes, int store) = 0;
};
extern thread_local ReplayObserver* g_replay_observer;

template <typename V>
__host__ __device__ inline void st(V* p, const V& v) {
#if defined(__HIP_DEVICE_COMPILE__)
#if LRT_NT_STORES
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
#else
  if (g_replay_observer) g_replay_observer->access(p, (int)sizeof(V), 1);
  *p = v;
#endif
}
template <typename V>
__host__ __device__ inline V ld(const V* p) {
#if defined(__HIP_DEVICE_COMPILE__)
#if LRT_NT_LOADS
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
#else
  if (g_replay_observer) g_replay_observer->access(p, (int)sizeof(V), 0);
  return *p;
#endif
}

struct TileCtx {
  int x0, y0;
  int off_in, off_out;
  bool full;
};

// ---------------------------------------------------------------------------
// Tiled: a TX x TY tile of the (x, y) plane goes through LDS. Reads walk x
// (contiguous in the input) with VX elements per lane; writes walk y
// (contiguous in the output) with VY elements per lane. One LDS row of pitch
// TX + 1 per y keeps the column reads of the write phase spread over banks.
// When x or y is a group of dims, each row (column) finds its offset on the
// other side from the group's digits; a full tile of one-dim groups steps
// through memory by a fixed stride instead.
// Reference counterpart: transposeTiled, src/kernel.cpp:48-170.
// ---------------------------------------------------------------------------
// G = false compiles only the one-dim-group path (the kernel then needs no
// registers for group maps); G = true handles groups of several dims too.
template <int B, int TX, int TY, int VX, int VY, bool G = false>
struct TiledOps {
  using E = typename Word<B>::type;
  using VI = typename Word<B * VX>::type;
  using VO = typename Word<B * VY>::type;
  static constexpr int NT = 256;
  static constexpr int LR = TX / VX;   // lanes per input row
  static constexpr int RP = NT / LR;   // input rows per pass
  static constexpr int NPR = TY / RP;  // read passes
  static constexpr int LW = TY / VY;   // lanes per output row
  static constexpr int WP = NT / LW;   // output rows (x) per pass
  static constexpr int NPW = TX / WP;  // write passes
  static constexpr int PITCH = TX + 1;
  // G: after the tile, TY input offsets of the tile's rows and TX output
  // offsets of its columns (ints), written once per tile by prepare().
  static constexpr int OFF_ELEMS = G ? ((TX + TY) * 4 + B - 1) / B : 0;
  static constexpr int LDS_ELEMS = TY * PITCH + OFF_ELEMS;
  static_assert(TX % VX == 0 && NT % LR == 0 && TY % RP == 0 && NPR >= 1, "read shape");
  static_assert(TY % VY == 0 && NT % LW == 0 && TX % WP == 0 && NPW >= 1, "write shape");
  static_assert(!G || TX + TY <= NT, "one thread per tile row and column");

  __host__ __device__ static int* offsets(E* lds) { return reinterpret_cast<int*>(lds + TY * PITCH); }
  __host__ __device__ static const int* offsets(const E* lds) {
    return reinterpret_cast<const int*>(lds + TY * PITCH);
  }

  // G only, before load(): thread t < TY finds the input offset of tile row
  // t, thread TY + i the output offset of tile column i (group digits, one
  // magic-number division each), so the load and store loops read them from
  // LDS instead of each thread dividing for every row it moves. A barrier
  // must separate prepare() from load().
  __host__ __device__ static void prepare(const TiledArgs& a, const TileCtx& c, int t, E* lds) {
    if constexpr (G) {
      int* o = offsets(lds);
      if (t < TY) {
        const int y = c.y0 + t;
        o[t] = y < a.Y ? tiled_in_y(a, (uint32_t)y) : 0;
      } else if (t < TY + TX) {
        const int x = c.x0 + (t - TY);
        o[t] = x < a.X ? tiled_out_x(a, (uint32_t)x) : 0;
      }
    }
  }

  __host__ __device__ static TileCtx item(const TiledArgs& a, uint32_t w) {
    TileCtx c;
    uint32_t mb = fdiv(w, a.ntiles_div);
    uint32_t tile = w - mb * (uint32_t)a.ntiles;
    uint32_t tx, ty;
    if (a.group <= 1) {
      ty = fdiv(tile, a.ntx_div);
      tx = tile - ty * (uint32_t)a.ntx;
    } else {
      const uint32_t g = fdiv(tile, a.group_row_div);  // group index
      const uint32_t within = tile - g * a.group_row_div.d;
      const FastDiv& rows = (tile >= (uint32_t)a.last_group_start) ? a.last_div : a.group_div;
      tx = fdiv(within, rows);
      ty = g * (uint32_t)a.group + (within - tx * rows.d);
    }
    c.x0 = (int)tx * TX;
    c.y0 = (int)ty * TY;
    outer_offsets(a.outer, mb, c.off_in, c.off_out);
    c.full = (c.x0 + TX <= a.X) && (c.y0 + TY <= a.Y);
    return c;
  }

  __host__ __device__ static void load(const TiledArgs& a, const TileCtx& c, int t,
                                       const E* __restrict__ in, E* lds) {
    VI r[NPR];
    fetch(a, c, t, in, r, lds);
    stash(a, c, t, r, lds);
  }

  // The two halves of load(): the tile's input rows into registers, then
  // registers into LDS (the pipelined kernel fetches the next tile between
  // them and the current tile's store phase).
  __host__ __device__ static void fetch(const TiledArgs& a, const TileCtx& c, int t, const E* __restrict__ in,
                                        VI* r, const E* lds) {
    const int lx = (t % LR) * VX;
    const int ly = t / LR;
    const int x = c.x0 + lx;
    if (c.full && (!G || a.yg.n == 1)) {
      const E* p = in + (c.off_in + x + (c.y0 + ly) * a.in_ys);
      const int step = RP * a.in_ys;
#pragma unroll
      for (int j = 0; j < NPR; j++) r[j] = ld(reinterpret_cast<const VI*>(p + j * step));
    } else {
      const bool xin = x < a.X;
#pragma unroll
      for (int j = 0; j < NPR; j++) {
        const int y = c.y0 + ly + j * RP;
        const int yoff = G ? offsets(lds)[ly + j * RP] : y * a.in_ys;
        if (xin && y < a.Y) r[j] = ld(reinterpret_cast<const VI*>(in + (c.off_in + x + yoff)));
      }
    }
  }
  __host__ __device__ static void stash(const TiledArgs& a, const TileCtx& c, int t, const VI* r, E* lds) {
    const int lx = (t % LR) * VX;
    const int ly = t / LR;
    const bool xin = c.x0 + lx < a.X;
#pragma unroll
    for (int j = 0; j < NPR; j++) {
      if (c.full || (xin && c.y0 + ly + j * RP < a.Y)) {
#pragma unroll
        for (int e = 0; e < VX; e++) lds[(ly + j * RP) * PITCH + lx + e] = vget<E>(r[j], e);
      }
    }
  }

  __host__ __device__ static void store(const TiledArgs& a, const TileCtx& c, int t,
                                        E* __restrict__ out, const E* lds) {
    const int ly = (t % LW) * VY;
    const int lx = t / LW;
    const int y = c.y0 + ly;
    VO r[NPW];
#pragma unroll
    for (int j = 0; j < NPW; j++)
#pragma unroll
      for (int e = 0; e < VY; e++) vset<E>(r[j], e, lds[(ly + e) * PITCH + lx + j * WP]);
    if (c.full && (!G || a.xg.n == 1)) {
      E* p = out + (c.off_out + y + (c.x0 + lx) * a.out_xs);
      const int step = WP * a.out_xs;
#pragma unroll
      for (int j = 0; j < NPW; j++) st(reinterpret_cast<VO*>(p + j * step), r[j]);
    } else {
      const bool yin = y < a.Y;
#pragma unroll
      for (int j = 0; j < NPW; j++) {
        const int x = c.x0 + lx + j * WP;
        const int xoff = G ? offsets(lds)[TY + lx + j * WP] : x * a.out_xs;
        if (yin && x < a.X) st(reinterpret_cast<VO*>(out + (c.off_out + y + xoff)), r[j]);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// TiledCopy: rows that are contiguous in both tensors, moved as W-byte words
// straight through registers. TX words per tile row, 256 / TX rows per pass,
// NP passes. Reference counterpart: transposeTiledCopy, src/kernel.cpp:453-545.
// ---------------------------------------------------------------------------
template <int W, int TX, int NP>
struct CopyOps {
  using V = typename Word<W>::type;
  static constexpr int NT = 256;
  static constexpr int RP = NT / TX;
  static constexpr int TY = RP * NP;
  static_assert(NT % TX == 0, "copy shape");

  __host__ __device__ static TileCtx item(const CopyArgs& a, uint32_t w) {
    TileCtx c;
    uint32_t mb = fdiv(w, a.ntiles_div);
    uint32_t tile = w - mb * (uint32_t)a.ntiles;
    uint32_t ty = fdiv(tile, a.ntx_div);
    uint32_t tx = tile - ty * (uint32_t)a.ntx;
    c.x0 = (int)tx * TX;
    c.y0 = (int)ty * TY;
    outer_offsets(a.outer, mb, c.off_in, c.off_out);
    c.full = (c.x0 + TX <= a.X) && (c.y0 + TY <= a.Y);
    return c;
  }

  __host__ __device__ static void run(const CopyArgs& a, const TileCtx& c, int t,
                                      const V* __restrict__ in, V* __restrict__ out) {
    const int lx = t % TX;
    const int ly = t / TX;
    const int x = c.x0 + lx;
    V r[NP];
    if (c.full) {
      const V* p = in + (c.off_in + x + (c.y0 + ly) * a.in_ys);
      const int step = RP * a.in_ys;
#pragma unroll
      for (int j = 0; j < NP; j++) r[j] = ld(p + j * step);
      V* q = out + (c.off_out + x + (c.y0 + ly) * a.out_ys);
      const int ostep = RP * a.out_ys;
#pragma unroll
      for (int j = 0; j < NP; j++) st(q + j * ostep, r[j]);
    } else {
      const bool xin = x < a.X;
#pragma unroll
      for (int j = 0; j < NP; j++) {
        const int y = c.y0 + ly + j * RP;
        if (xin && y < a.Y) r[j] = ld(in + (c.off_in + x + y * a.in_ys));
      }
#pragma unroll
      for (int j = 0; j < NP; j++) {
        const int y = c.y0 + ly + j * RP;
        if (xin && y < a.Y) st(out + (c.off_out + x + y * a.out_ys), r[j]);
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Packed / PackedSplit: a whole on-chip sub-tensor (one chunk of the split dim
// for PackedSplit) goes through LDS per outer index. Each thread owns NREG
// positions of the staged volume and keeps their input offset, output offset
// and LDS read position in registers for all outer indices of its block.
// Reference counterparts: transposePacked (src/kernel.cpp:175-296) and
// transposePackedSplit (src/kernel.cpp:304-444).
// ---------------------------------------------------------------------------
template <int B, int NREG>
struct PackedOps {
  using E = typename Word<B>::type;

  struct Pre {
    int in[NREG];
    int out[NREG];
    int sh[NREG];
    int vol;
    int base_in, base_out;
    FastDiv pad;
  };

  __host__ __device__ static void prepare(const PackedArgs& a, int split, int t, int nt, Pre& p) {
    const int q0 = a.split_dim / a.num_split;
    const int c0 = (int)((long long)split * a.split_dim / a.num_split);
    const int c1 = (int)((long long)(split + 1) * a.split_dim / a.num_split);
    const int plus = (c1 - c0) - q0;
    const StagedShape& s = a.shape[plus];
    p.vol = s.vol;
    p.pad = s.pad;
    p.base_in = c0 * a.split_in_stride;
    p.base_out = c0 * a.split_out_stride;
#pragma unroll
    for (int r = 0; r < NREG; r++) {
      uint32_t k = (uint32_t)(t + r * nt);
      int oi = 0;
      for (int d = 0; d < s.n; d++) {
        uint32_t q = fdiv(k, s.din[d]);
        oi += (int)(k - q * s.din[d].d) * s.ct_in[d];
        k = q;
      }
      k = (uint32_t)(t + r * nt);
      int oo = 0, sh = 0;
      for (int d = 0; d < s.n; d++) {
        uint32_t q = fdiv(k, s.dout[d]);
        int digit = (int)(k - q * s.dout[d].d);
        oo += digit * s.ct_out[d];
        sh += digit * s.ct_sh[d];
        k = q;
      }
      p.in[r] = oi;
      p.out[r] = oo;
      p.sh[r] = sh;
    }
  }

  __host__ __device__ static void load(const Pre& p, int off_in, int t, int nt,
                                       const E* __restrict__ in, E* lds) {
    E r[NREG];
    fetch(p, off_in, t, nt, in, r);
    stash(p, t, nt, r, lds);
  }

  // The two halves of load(): global memory to registers, registers to LDS.
  __host__ __device__ static void fetch(const Pre& p, int off_in, int t, int nt, const E* __restrict__ in,
                                        E* r) {
    const E* base = in + (p.base_in + off_in);
#pragma unroll
    for (int j = 0; j < NREG; j++)
      if (t + j * nt < p.vol) r[j] = ld(base + p.in[j]);
  }
  __host__ __device__ static void stash(const Pre& p, int t, int nt, const E* r, E* lds) {
#pragma unroll
    for (int j = 0; j < NREG; j++)
      if (t + j * nt < p.vol) {
        const uint32_t q = (uint32_t)(t + j * nt);
        lds[q + fdiv(q, p.pad)] = r[j];
      }
  }

  __host__ __device__ static void store(const Pre& p, int off_out, int t, int nt,
                                        E* __restrict__ out, const E* lds) {
    E r[NREG];
#pragma unroll
    for (int j = 0; j < NREG; j++)
      if (t + j * nt < p.vol) r[j] = lds[p.sh[j]];
    E* base = out + (p.base_out + off_out);
#pragma unroll
    for (int j = 0; j < NREG; j++)
      if (t + j * nt < p.vol) st(base + p.out[j], r[j]);
  }
};

}  // namespace lrt
