// common.hpp -- shared device-side definitions of the gfx950 point-to-sparse-grid library.
//
// The topology is the reference's pool layout (src/gvdb_node.h:27-40, A.h:59-62): per level a
// pool of 64-byte nodes and, for levels >= 1, a pool of child lists of res^3 uint64 entries.
// Child list i belongs to node i of the same level (both pools grow in lockstep, V.cpp:1308-1331),
// and every node is aligned to its level's range, so a tree descent needs one 8-byte load per
// level and never reads a node (the reference reads node + child list per level, N.cuh:255-288).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gvdbx.h"

namespace gvdbx {

constexpr int kMaxLev = 10;
constexpr uint64_t kUndef64 = 0xFFFFFFFFFFFFFFFFull;  // ID_UNDEF64
constexpr uint64_t kUndefL = 0xFFFFFFFFull;           // ID_UNDEFL
constexpr float kNoHit = 1.0e10f;                     // NOHIT (src/gvdb_types.h:52)

// Tree description passed by value to kernels.
struct DevTree {
  int depth;                 // root level
  int levs;
  int ld[kMaxLev];           // log2 of per-level resolution
  int lgR[kMaxLev];          // log2 of per-level range in voxels
  int root[3];               // root corner (aligned to its range)
  gvdbx_node* nodes[kMaxLev];
  uint64_t* clist[kMaxLev];
  // Leaf lookup table over a box of leaf-brick cells inside the root, or null: cell
  // (bx, by, bz) (brick coordinates, global), c = ((bz - lo) * n1 + by - lo) * n0 + bx - lo,
  // holds at bmap[2c] the leaf whose brick it is and at bmap[2c + 1] that leaf's atlas slot
  // (both -1 where none is linked), so one 8-byte load takes a point to its slot.  Built from
  // the slots, which are exactly the linked leaves, so a lookup inside the box equals the descent.
  const int* bmap;
  int bm_lo[3], bm_n[3];
};

// A dispatch carries at most 2^32 - 1 work-items per grid dimension (the grid sizes of
// hsa_kernel_dispatch_packet_t are 32-bit), so a kernel that gives one workgroup to each brick,
// slot or subcell would overflow it from 2^32 / blockDim items on (16.7M bricks at 256 threads).
// Such kernels lay their workgroups out on a 2D grid of at most kGridCols columns, recover the
// linear item with item_index() and return when it is past the item count.
constexpr int64_t kGridCols = 65536;
inline dim3 item_grid(int64_t items) {
  if (items < 1) items = 1;
  const int64_t cols = items < kGridCols ? items : kGridCols;
  return dim3((unsigned)cols, (unsigned)((items + cols - 1) / cols));
}
__device__ __forceinline__ int64_t item_index() { return (int64_t)blockIdx.y * gridDim.x + blockIdx.x; }

// XCD-aware items: workgroups are dispatched to the 8 XCDs round-robin, so workgroup w runs on
// XCD w % 8.  Giving XCD k the contiguous items [k * chunk, (k + 1) * chunk) in order keeps items
// that are neighbours in the item order (bricks in slot order) on one XCD, where they share its
// L2.  The grid holds xcd_items(n) items; the caller skips those at or past n.
constexpr int kXcds = 8;
inline int64_t xcd_items(int64_t n) { return ((n + kXcds - 1) / kXcds) * kXcds; }
__device__ __forceinline__ int64_t xcd_item(int64_t n) {
  const int64_t w = item_index();
  const int64_t chunk = (n + kXcds - 1) / kXcds;
  if (w >= kXcds * chunk) return n;  // the surplus of a grid rounded up further
  return (w % kXcds) * chunk + w / kXcds;
}

// Launch shape of the per-brick kernels (count, insert, density gather).
struct LaunchCfg {
  int block;    // threads per brick workgroup (64, 128 or 256); 0 = chosen from the items per brick
  int persist;  // 1: a grid of resident workgroups takes bricks from a queue
  int* queue;   // the queue counter (one device int), cleared before every persistent launch
  int cus;      // compute units of the device (the persistent grid is a multiple of it)
};

// The bricks of a per-brick kernel: one per workgroup (grid form: m = the item index, or with
// xcd xcd_item, so that bricks that are neighbours in slot order, whose points share cache lines
// of the caller's arrays, run on one XCD; the grid holds xcd_items(n) items), or, persistent,
// taken from the queue kBrickChunk at a time until it passes n.  The body may return
// early only uniformly over the workgroup; a barrier follows every brick, so the body may reuse
// its LDS from one brick to the next.  Every workgroup leaves once the queue passes n.
constexpr int kBrickChunk = 4;
template <bool kPersist, typename F>
__device__ __forceinline__ void for_each_brick(int64_t n, int* queue, F&& body, bool xcd = false) {
  if (!kPersist) {
    const int64_t m = xcd ? xcd_item(n) : item_index();
    if (m < n) body((int)m);
    return;
  }
  __shared__ int first;
  for (;;) {
    if (threadIdx.x == 0) first = atomicAdd(queue, kBrickChunk);
    __syncthreads();
    const int f = first;
    __syncthreads();
    if ((int64_t)f >= n) return;
    const int last = (int64_t)f + kBrickChunk < n ? f + kBrickChunk : (int)n;
    for (int m = f; m < last; ++m) {
      body(m);
      __syncthreads();
    }
  }
}

__device__ __forceinline__ uint64_t elem0(int lev, uint64_t ndx) { return ((uint64_t)lev << 8) | (ndx << 16); }

__device__ __forceinline__ bool in_root(const DevTree& T, int x, int y, int z) {
  const int R = T.lgR[T.depth];
  return x >= T.root[0] && y >= T.root[1] && z >= T.root[2] && ((x - T.root[0]) >> R) == 0 &&
         ((y - T.root[1]) >> R) == 0 && ((z - T.root[2]) >> R) == 0;
}

// Child slot of position (x,y,z) inside its level-l ancestor: ((z*res + y)*res + x) of the
// child cell (getBitPos, N.cuh:172-176).
__device__ __forceinline__ int child_bit(const DevTree& T, int l, int x, int y, int z) {
  const int sh = T.lgR[l - 1], ld = T.ld[l], m = (1 << ld) - 1;
  return ((((z >> sh) & m) << ld | ((y >> sh) & m)) << ld) | ((x >> sh) & m);
}

// Index of the linked node of level stop containing (x,y,z), or -1 (positions outside the
// root are missing).  Restates getPosParent/getPosLeafParent (N.cuh:255-348).
__device__ __forceinline__ int64_t descend(const DevTree& T, int x, int y, int z, int stop) {
  if (!in_root(T, x, y, z)) return -1;
  int64_t cur = 0;
  for (int l = T.depth; l > stop; --l) {
    const uint64_t e = T.clist[l][((uint64_t)cur << (3 * T.ld[l])) + (uint64_t)child_bit(T, l, x, y, z)];
    if (e == kUndef64) return -1;
    cur = (int64_t)(e >> 16);
  }
  return cur;
}

// The lookup table's (leaf, slot) for voxel (x, y, z): (-1, -1) when no leaf is linked there,
// (-2, -2) when there is no table or the voxel's brick lies outside its box.
__device__ __forceinline__ int2 table_entry(const DevTree& T, int x, int y, int z) {
  if (!T.bmap) return make_int2(-2, -2);
  const int s = T.lgR[0];
  const unsigned bx = (unsigned)((x >> s) - T.bm_lo[0]), by = (unsigned)((y >> s) - T.bm_lo[1]),
                 bz = (unsigned)((z >> s) - T.bm_lo[2]);
  if (bx >= (unsigned)T.bm_n[0] || by >= (unsigned)T.bm_n[1] || bz >= (unsigned)T.bm_n[2]) return make_int2(-2, -2);
  return reinterpret_cast<const int2*>(T.bmap)[((int64_t)bz * T.bm_n[1] + by) * T.bm_n[0] + bx];
}

// The lookup table's leaf for voxel (x, y, z): the leaf index, -1 when no leaf is linked there,
// -2 when there is no table or the voxel's brick lies outside its box.
__device__ __forceinline__ int64_t table_leaf(const DevTree& T, int x, int y, int z) {
  return table_entry(T, x, y, z).x;
}

// The leaf holding voxel (x, y, z), or -1: descend(T, x, y, z, 0), through the lookup table
// when the voxel's brick lies in its box.
__device__ __forceinline__ int64_t find_leaf(const DevTree& T, int x, int y, int z) {
  const int64_t lf = table_leaf(T, x, y, z);
  return lf != -2 ? lf : descend(T, x, y, z, 0);
}

// The atlas slot of the leaf holding voxel (x, y, z), or -1 (no leaf, or a leaf without slot):
// one table load inside the table's box, else the descent and leaf_slot.
__device__ __forceinline__ int find_slot(const DevTree& T, int x, int y, int z, const int* __restrict__ leaf_slot) {
  const int2 e = table_entry(T, x, y, z);
  if (e.x != -2) return e.y;
  const int64_t lf = descend(T, x, y, z, 0);
  return lf < 0 ? -1 : leaf_slot[lf];
}

// |d|^2 with a fixed operation order, so every kernel that reduces distances (gathers,
// scatter-reduce) rounds identically.
__device__ __forceinline__ float dist2(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

// x / 2^s * 2^s with C's truncation toward zero (the reference's integer division of a voxel
// coordinate by the leaf range or the subcell size), by shifts: negative x is biased by 2^s - 1
// first
__device__ __forceinline__ int trunc_multiple(int x, int s) { return ((x + ((x >> 31) & ((1 << s) - 1))) >> s) << s; }

// 16-byte store marked non-temporal (the line is not kept in the caches): for output written
// once and not read again by the kernels that follow
typedef float f32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_stream4(float* p, float a, float b, float c, float d) {
  const f32x4_t v = {a, b, c, d};
  __builtin_nontemporal_store(v, reinterpret_cast<f32x4_t*>(p));
}

// ---- wave64 helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Exclusive scan of a 256-thread block's values; returns the block total in *total.
__device__ __forceinline__ int block_exclusive_scan_256(int v, int* total) {
  __shared__ int ws[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[wave] = x;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    int s = ws[w];
    pre += (w < wave) ? s : 0;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return pre + x - v;
}

// Monotonic float <-> int mapping for atomic min/max on floats.
__device__ __forceinline__ int float_to_ordered(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__host__ __device__ inline float ordered_to_float(int i) {
  int b = i >= 0 ? i : i ^ 0x7FFFFFFF;
  float f;
  __builtin_memcpy(&f, &b, 4);
  return f;
}

__device__ __forceinline__ float3 load_f3(const char* base, int64_t i, int off, int stride) {
  const float* p = (const float*)(base + i * (int64_t)stride + off);
  return make_float3(p[0], p[1], p[2]);
}

// ushort((v - lo) / range * 65535) of gvdbInsertSubcell_fp16 (P.cuh:200-207), in IEEE single
// operations in that order, saturated to [0, 65535] (NaN -> 0)
__device__ __forceinline__ uint16_t quant16(float v, float lo, float range) {
  const float f = __fmul_rn(__fdiv_rn(__fsub_rn(v, lo), range), 65535.0f);
  return (uint16_t)(f >= 65535.0f ? 65535u : (f >= 0.0f ? (uint32_t)f : 0u));
}
// u / 65535 * range + lo of gvdbGatherLevelSet_fp16 (P.cuh:791-793), unfused
__device__ __forceinline__ float dequant16(uint16_t u, float lo, float range) {
  return __fadd_rn(__fmul_rn(__fdiv_rn((float)u, 65535.0f), range), lo);
}

__device__ __forceinline__ uint32_t load_u32(const char* base, int64_t i, int off, int stride) {
  return *(const uint32_t*)(base + i * (int64_t)stride + off);
}

}  // namespace gvdbx
