// Shared host/device definitions for the gfx950 transpose library.
//
// Everything here is plain data passed by value in kernel arguments, so a
// plan needs no device-side tables (the reference uploads TensorConv tables
// with hipMalloc + hipMemcpyAsync in librettPlan_t::activate,
// src/plan.cpp:1648-1676; here the same information travels in the kernel
// argument segment and is read with scalar loads).
#pragma once

#include <stdint.h>
#include <hip/hip_runtime.h>

namespace lrt {

// Largest rank after size-1 dims are removed. Every remaining dim is >= 2 and
// the volume is < 2^31, so at most 30 dims remain.
constexpr int kMaxRank = 32;
// Largest number of dims staged on chip by the Packed kernels.
constexpr int kMaxMmk = 12;

// Method numbering follows the reference enum (src/plan.h:54-56).
enum Method : int {
  kUnknown = 0,
  kTrivial = 1,
  kPacked = 2,
  kPackedSplit = 3,
  kTiled = 4,
  kTiledCopy = 5,
  kNumMethods = 6
};

// Division by an invariant 32-bit divisor for dividends below 2^31
// (Granlund and Montgomery, "Division by invariant integers using
// multiplication", PLDI 1994, round-up variant): with l = ceil(log2 d) and
// m = floor(2^32 (2^l - d) / d) + 1, n / d = (umulhi(n, m) + n) >> l.
// umulhi(n, m) <= n, so the sum stays below 2^32 while n < 2^31.
struct FastDiv {
  uint32_t mul;
  uint32_t shr;
  uint32_t d;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  uint64_t m = ((1ull << 32) * ((1ull << l) - d)) / d + 1;
  f.mul = (uint32_t)m;
  f.shr = l;
  f.d = d;
  return f;
}

__host__ __device__ inline uint32_t umulhi32(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umulhi(a, b);
#else
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
#endif
}

__host__ __device__ inline uint32_t fdiv(uint32_t n, const FastDiv& f) {
  return (umulhi32(n, f.mul) + n) >> f.shr;
}

// Outer ("Mbar") index space: dims that are neither staged on chip nor part
// of the tile plane, in input order. Index k of this space has digits
// (k % d0, (k / d0) % d1, ...); digit r moves the input by ct_in[r] and the
// output by ct_out[r] elements (the reference keeps the same information as
// TensorConvInOut entries, src/plan.cpp:935-947).
struct OuterMap {
  int n;
  FastDiv div[kMaxRank];
  int ct_in[kMaxRank];
  int ct_out[kMaxRank];
};

__host__ __device__ inline void outer_offsets(const OuterMap& m, uint32_t k,
                                              int& off_in, int& off_out) {
  int oi = 0, oo = 0;
  for (int r = 0; r < m.n; r++) {
    uint32_t q = fdiv(k, m.div[r]);
    int digit = (int)(k - q * m.div[r].d);
    oi += digit * m.ct_in[r];
    oo += digit * m.ct_out[r];
    k = q;
  }
  off_in = oi;
  off_out = oo;
}

// A group of dims walked as one linear index v of the tile plane: digit k of
// v (extent of dim k, the first dim fastest) moves the other tensor by
// stride[k] elements. A group of one dim is linear (n == 1: v * stride[0]).
constexpr int kMaxGroup = 6;
struct GroupMap {
  int n;
  FastDiv div[kMaxGroup - 1];  // extents of all but the last dim
  int stride[kMaxGroup];
};

__host__ __device__ inline int group_offset(const GroupMap& g, uint32_t v) {
  int off = 0;
#pragma unroll
  for (int k = 0; k < kMaxGroup - 1; k++) {
    if (k + 1 >= g.n) return off + (int)v * g.stride[k];
    const uint32_t q = fdiv(v, g.div[k]);
    off += (int)(v - q * g.div[k].d) * g.stride[k];
    v = q;
  }
  return off + (int)v * g.stride[kMaxGroup - 1];
}

// Tiled: 2-D tile plane of (x = input dim 0, y = the input dim that becomes
// output dim 0). Reference counterpart: transposeTiled, src/kernel.cpp:48-170.
// x may also be a group of leading input dims (contiguous in the input, in
// any order in the output) and y a group of leading output dims (contiguous
// in the output, in any order in the input), so that tensors whose first
// dims are short still get tiles with long runs on both sides; the output
// offset of x (the input offset of y) is then group_offset(xg, x)
// (group_offset(yg, y)) instead of x * out_xs (y * in_ys).
struct TiledArgs {
  int X, Y;          // plane extents in elements
  int in_ys;         // input stride of y (yg.n == 1)
  int out_xs;        // output stride of x (xg.n == 1)
  GroupMap xg;       // x digits -> output offset
  GroupMap yg;       // y digits -> input offset
  int ntx;           // tiles along x
  int ntiles;        // tiles per plane
  FastDiv ntx_div;
  FastDiv ntiles_div;
  int items;         // ntiles * volMbar
  // Tile order: tiles go in groups of `group` tile rows (along y); within a
  // group, y varies fastest, so blocks dispatched together write adjacent
  // stretches of the same output rows as well as reading adjacent stretches
  // of the same input rows.
  int group;
  FastDiv group_div;      // group
  FastDiv group_row_div;  // group * ntx
  FastDiv last_div;       // rows in the last (possibly partial) group
  int last_group_start;   // first tile index of the last group
  // > 1: workgroup b starts at tile xcd_logical(b, grid, xcds), so each XCD
  // walks one contiguous range of tiles (tiles that share 128-byte lines
  // then meet in one XCD's L2); 1: workgroup b starts at tile b.
  int xcds;
  OuterMap outer;
};

// Output offset of tile-plane column x and input offset of row y.
__host__ __device__ inline int tiled_out_x(const TiledArgs& a, uint32_t x) {
  return a.xg.n > 1 ? group_offset(a.xg, x) : (int)x * a.out_xs;
}
__host__ __device__ inline int tiled_in_y(const TiledArgs& a, uint32_t y) {
  return a.yg.n > 1 ? group_offset(a.yg, y) : (int)y * a.in_ys;
}

// TiledCopy: rows that stay contiguous in input and output (perm[0] == 0),
// moved as W-byte words. Plane = (x = words of dim 0, y = the input dim that
// becomes output dim 1). Reference counterpart: transposeTiledCopy,
// src/kernel.cpp:453-545.
struct CopyArgs {
  int X, Y;          // plane extents (x in words)
  int in_ys;         // input stride of y in words
  int out_ys;        // output stride of y in words
  int ntx;
  int ntiles;
  FastDiv ntx_div;
  FastDiv ntiles_div;
  int items;
  int xcds;          // work order across XCDs, as in TiledArgs
  OuterMap outer;    // strides in words
};

// One on-chip ("Mmk") sub-tensor shape. The LDS holds it in input order, so
// LDS position p is also its input-order linear index. Input-order digits of
// p give the global input offset; output-order digits of the output-ordered
// index u give the global output offset and the LDS position to read.
// Reference counterpart: TensorConvInOut Mmk and TensorConv Msh tables,
// src/plan.cpp:1009-1058 (Packed) and :953-1007 (PackedSplit).
struct StagedShape {
  int vol;
  int n;
  // LDS padding: position p is stored at p + p / pad.d. pad.d is the LDS
  // stride S of output dim 0 when S is even, so that the write phase, whose
  // consecutive lanes step by S in LDS, steps by the odd S + 1 instead and
  // touches distinct banks. ct_sh already includes the padding.
  FastDiv pad;
  FastDiv din[kMaxMmk];
  int ct_in[kMaxMmk];
  FastDiv dout[kMaxMmk];
  int ct_out[kMaxMmk];
  int ct_sh[kMaxMmk];
};

// Packed and PackedSplit share one kernel: Packed is a split count of 1.
// Reference counterparts: transposePacked (src/kernel.cpp:175-296) and
// transposePackedSplit (src/kernel.cpp:304-444).
struct PackedArgs {
  StagedShape shape[2];  // [0]: chunk of floor(splitDim/numSplit), [1]: +1
  int num_split;
  int split_dim;
  int split_in_stride;   // input stride of the split dim
  int split_out_stride;  // output stride of the split dim
  int vol_outer;
  int outer_per_block;   // order 0: contiguous outer indices per block row
  // Work order. 0: block (x, y) takes chunk x and outer indices
  // [y * outer_per_block, (y + 1) * outer_per_block). 1 (XCD-interleaved):
  // the linear block id is remapped so that the blocks one XCD runs at the
  // same time take consecutive outer indices of one chunk; a block then
  // strides by gridDim.y. Adjacent outer indices that share cache lines are
  // then fetched once into that XCD's L2.
  int order;
  int xcds;              // XCDs the dispatcher deals workgroups to, round-robin
  OuterMap outer;
};

// Order-1 mapping: linear block b of G -> logical block; blocks b, b + x,
// b + 2x, ... (one XCD, dispatched together) get consecutive logical ids.
__host__ __device__ inline int xcd_logical(int b, int G, int xcds) {
  const int q = G / xcds, r = G - (G / xcds) * xcds;
  const int x = b - (b / xcds) * xcds;
  return x * q + (x < r ? x : r) + b / xcds;
}

}  // namespace lrt
