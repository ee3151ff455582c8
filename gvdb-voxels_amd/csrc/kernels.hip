// kernels.hip -- gfx950 kernels of the point-to-sparse-grid path.
//
// Each kernel cites the reference kernel whose semantics it restates
// (P.cuh = kernels/cuda_gvdb_particles.cuh, N.cuh = kernels/cuda_gvdb_nodes.cuh,
//  O.cuh = kernels/cuda_gvdb_operators.cuh, R.cuh = kernels/cuda_gvdb_raycast.cuh).
#include <limits.h>
#include <string.h>

#include "kernels.hpp"
#include "primitives.hpp"

namespace gvdbx {

static inline int blocks_for(int64_t n, int bs, int64_t cap = 0x7FFFFFFF) {
  int64_t g = (n + bs - 1) / bs;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// ================================================================ keys

template <typename K>
__device__ __forceinline__ K pack_key(KeyFmt f, int lev, int xr, int yr, int zr) {
  return ((K)lev << (3 * f.b)) | ((K)xr << (2 * f.b)) | ((K)yr << f.b) | (K)zr;
}

// gvdbCalcBrickId (P.cuh:352-382): per point and level < nlev the covering node of q (nlev =
// depth gives every level; the full pass takes the leaf level only and derives the rest from
// the unique leaves, k_expand_ancestors).  For q >= 0, floor(q / R) (the ceilf-and-adjust form
// of GetCoveringNode, P.cuh:88-100, with R a power of two) equals trunc(q) >> log2(R).
template <typename K>
__global__ void __launch_bounds__(256) k_brick_keys(DevTree T, PointsView P, KeyFmt f, int nlev,
                                                   K* __restrict__ keys) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const int d = T.depth;
  float3 p = load_f3(P.pos, i, P.pos_off, P.pos_stride);
  const float q[3] = {p.x + P.tx, p.y + P.ty, p.z + P.tz};
  K* out = keys + i * nlev;
  if (q[0] < 0 || q[1] < 0 || q[2] < 0) {
    for (int l = 0; l < nlev; ++l) out[l] = pack_key<K>(f, d, 0, 0, 0);
    return;
  }
  const int3 o = cloud_origin(P, i);
  const int iq[3] = {(int)q[0] + o.x, (int)q[1] + o.y, (int)q[2] + o.z};
  for (int l = 0; l < nlev; ++l) {
    const int sh = T.lgR[l];
    out[l] = pack_key<K>(f, l, (iq[0] >> sh) - (T.root[0] >> sh), (iq[1] >> sh) - (T.root[1] >> sh),
                         (iq[2] >> sh) - (T.root[2] >> sh));
  }
}

template <typename K>
void brick_keys(const DevTree& T, const PointsView& P, KeyFmt f, int nlev, K* keys, hipStream_t s) {
  hipLaunchKernelGGL(k_brick_keys<K>, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, T, P, f, nlev, keys);
}

// Every unique leaf key (flags / rank of the sorted leaf keys) expands into its own key and the
// keys of its ancestors at levels 1..depth-1, at out[rank * depth + level].  Root-relative
// coordinates shift down level by level, since the root corner is aligned to its range:
// (a >> s) - (root >> s) == ((a >> t) - (root >> t)) >> (s - t) for t <= s.
template <typename K>
__global__ void k_expand_ancestors(DevTree T, const K* __restrict__ sorted, const int* __restrict__ flags,
                                   const int* __restrict__ rank, int64_t n, KeyFmt f, K* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flags[i]) return;
  const int d = T.depth;
  const K v = sorted[i];
  const uint64_t mask = (1ull << f.b) - 1ull;
  const int xr = (int)((v >> (2 * f.b)) & mask), yr = (int)((v >> f.b) & mask), zr = (int)(v & mask);
  K* o = out + (int64_t)rank[i] * d;
  for (int l = 0; l < d; ++l) {
    const int sh = T.lgR[l] - T.lgR[0];
    o[l] = pack_key<K>(f, l, xr >> sh, yr >> sh, zr >> sh);
  }
}

template <typename K>
void expand_ancestors(const DevTree& T, const K* sorted, const int* flags, const int* rank, int64_t n, KeyFmt f,
                      K* out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_expand_ancestors<K>, dim3(blocks_for(n, 256)), dim3(256), 0, s, T, sorted, flags, rank, n, f,
                     out);
}

// Box enumeration of gvdbCalcExtraBrickId (P.cuh:467-512, hi = trunc(q) + 2r) and
// gvdbCalcIncreExtraBrickId (P.cuh:420-465, hi = trunc(q) + r, marks found leaves).
// The reference walks the box once per level and descends for every node.  The level-l box
// nodes are exactly the level-l ancestors of the leaf-level box nodes, so here each leaf-level
// node is descended once: a found leaf is marked, and when the descent stops at level l the
// missing chain (levels l-1 .. 0) is emitted (duplicates vanish in the unique step).
// Nodes outside the root are not emitted.
// Emission needs no device-scope atomic: each workgroup ranks its keys with an LDS counter into
// its own slice of `region` (cap_b = 256 * the per-point bound), writes its count, and a scan
// plus k_compact_regions pack the slices.
template <typename K, bool kTwo>
__global__ void __launch_bounds__(256) k_box_keys(DevTree T, PointsView P, KeyFmt f, float r, float hi_mul,
                                                  K* __restrict__ region, int64_t cap_b, int* __restrict__ blockcnt,
                                                  int* __restrict__ markers, K* __restrict__ compact, int compact_cap,
                                                  int* __restrict__ counter) {
  __shared__ int bcount;
  // kTwo: the box-node list of the workgroup (thread << 3 | k) and each thread's box corner
  __shared__ uint16_t s_list[kTwo ? 256 * 8 : 1];
  __shared__ int s_lo[kTwo ? 256 : 1][3];
  if (threadIdx.x == 0) bcount = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  K* out = region + (int64_t)blockIdx.x * cap_b;
  const int R0 = 1 << T.lgR[0];
  const int d = T.depth;
  auto emit = [&](int x, int y, int z, int64_t cur, int miss) {
    if (miss < 0) {
      // a plain store of the one value every writer stores: no read to wait for
      if (markers) markers[cur] = 1;
      return;
    }
    const int k = atomicAdd(&bcount, miss + 1);
    for (int l = miss; l >= 0; --l) {
      const int sh = T.lgR[l];
      out[k + miss - l] = pack_key<K>(f, l, (x >> sh) - (T.root[0] >> sh), (y >> sh) - (T.root[1] >> sh),
                                      (z >> sh) - (T.root[2] >> sh));
    }
  };
  // the leaf of box node (x, y, z) from the lookup table, else by descent; then emit
  auto visit = [&](int x, int y, int z) {
    const int64_t lf = table_leaf(T, x, y, z);
    if (lf >= 0) {  // found in the lookup table
      emit(x, y, z, lf, -1);
      return;
    }
    int64_t cur = 0;
    int miss = -1;  // highest missing level along the chain
    for (int l = d; l > 0; --l) {
      const uint64_t e = T.clist[l][((uint64_t)cur << (3 * T.ld[l])) + (uint64_t)child_bit(T, l, x, y, z)];
      if (e == kUndef64) {
        miss = l - 1;
        break;
      }
      cur = (int64_t)(e >> 16);
    }
    emit(x, y, z, cur, miss);
  };
  unsigned box_mask = 0;
  float q[3] = {-1.f, -1.f, -1.f};
  if (i < P.n) {
    const float3 p = load_f3(P.pos, i, P.pos_off, P.pos_stride);
    q[0] = p.x + P.tx;
    q[1] = p.y + P.ty;
    q[2] = p.z + P.tz;
  }
  if (q[0] >= 0 && q[1] >= 0 && q[2] >= 0) {
    int lo[3], hi[3];
    const int3 o = i < P.n ? cloud_origin(P, i) : make_int3(0, 0, 0);
    const int org[3] = {o.x, o.y, o.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const float tq = (float)(int)q[a];
      lo[a] = trunc_multiple((int)(tq - r), T.lgR[0]) + org[a];
      hi[a] = trunc_multiple((int)(tq + hi_mul * r), T.lgR[0]) + org[a];
    }
    if (kTwo) {
      // At most two leaf nodes per axis.  The point's box nodes (bit k = (jx, jy, jz) of mask:
      // corner lo + j R0 per axis, inside the box and the root) are listed after those of the
      // threads before it, and the workgroup then walks that list 256 entries at a time, so the
      // lookups of a round are independent and no load slot goes to a corner outside the box.
      const int Rr = T.lgR[T.depth];
      bool ok[3][2];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int c = lo[a] + j * R0;
          ok[a][j] = c <= hi[a] && c >= T.root[a] && ((c - T.root[a]) >> Rr) == 0;
        }
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (ok[0][k & 1] && ok[1][(k >> 1) & 1] && ok[2][k >> 2]) box_mask |= 1u << k;
#pragma unroll
      for (int a = 0; a < 3; ++a) s_lo[threadIdx.x][a] = lo[a];
    } else {
      for (int x = lo[0]; x <= hi[0]; x += R0)
        for (int y = lo[1]; y <= hi[1]; y += R0)
          for (int z = lo[2]; z <= hi[2]; z += R0)
            if (in_root(T, x, y, z)) visit(x, y, z);
    }
  }
  if (kTwo) {
    int total;
    int at = block_exclusive_scan_256(__popc(box_mask), &total);
    for (unsigned m = box_mask; m; m &= m - 1) s_list[at++] = (uint16_t)(threadIdx.x << 3 | (__ffs(m) - 1));
    __syncthreads();
    for (int base = 0; base < total; base += 256) {
      const int t = base + (int)threadIdx.x;
      if (t >= total) continue;
      const int e = s_list[t], src = e >> 3, k = e & 7;
      visit(s_lo[src][0] + (k & 1) * R0, s_lo[src][1] + ((k >> 1) & 1) * R0, s_lo[src][2] + (k >> 2) * R0);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) blockcnt[blockIdx.x] = bcount;
  if (compact) {
    // the keys also go to one packed array: a workgroup that found any claims its range with one
    // device atomic (few do, the keys of an incremental step come from the points near new
    // bricks), in no fixed order, which the sort that follows removes
    __shared__ int base;
    if (threadIdx.x == 0) base = bcount > 0 ? atomicAdd(counter, bcount) : 0;
    __syncthreads();
    for (int j = threadIdx.x; j < bcount; j += blockDim.x)
      if (base + j < compact_cap) compact[base + j] = out[j];
  }
}

template <typename K>
void box_keys(const DevTree& T, const PointsView& P, KeyFmt f, float r, float hi_mul, K* region, int64_t cap_b,
              int* blockcnt, int* markers, hipStream_t s, K* compact, int compact_cap, int* counter) {
  if (P.n <= 0) return;
  // a box [trunc(q) - r, trunc(q) + hi_mul r] spans at most 2 leaf nodes per axis when it is
  // shorter than one leaf node
  const bool two = (1.0 + (double)hi_mul) * (double)r + 1.0 < (double)(1 << T.lgR[0]);
  if (two)
    hipLaunchKernelGGL((k_box_keys<K, true>), dim3(blocks_for(P.n, 256)), dim3(256), 0, s, T, P, f, r, hi_mul, region,
                       cap_b, blockcnt, markers, compact, compact_cap, counter);
  else
    hipLaunchKernelGGL((k_box_keys<K, false>), dim3(blocks_for(P.n, 256)), dim3(256), 0, s, T, P, f, r, hi_mul,
                       region, cap_b, blockcnt, markers, compact, compact_cap, counter);
}

template <typename K>
__global__ void __launch_bounds__(256) k_compact_regions(const K* __restrict__ region, int64_t cap_b,
                                                         const int* __restrict__ blockcnt,
                                                         const int* __restrict__ blockoff, K* __restrict__ out,
                                                         int64_t out_cap) {
  const int64_t b = blockIdx.x;
  const int n = blockcnt[b];
  const int64_t o = blockoff[b];
  for (int j = threadIdx.x; j < n; j += blockDim.x)
    if (o + j < out_cap) out[o + j] = region[b * cap_b + j];
}

template <typename K>
void compact_regions(const K* region, int64_t cap_b, const int* blockcnt, const int* blockoff, int64_t nblocks, K* out,
                     int64_t out_cap, hipStream_t s) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(k_compact_regions<K>, dim3((unsigned)nblocks), dim3(256), 0, s, region, cap_b, blockcnt,
                     blockoff, out, out_cap);
}

// gvdbFindUnique (P.cuh:336-350): run starts of the sorted keys, invalid level excluded.
template <typename K>
__global__ void k_unique_flags(const K* __restrict__ k, int64_t m, KeyFmt f, int depth, int* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const K v = k[i];
  const int lev = (int)(v >> (3 * f.b));
  flags[i] = ((i == 0 || k[i - 1] != v) && lev < depth) ? 1 : 0;
}

template <typename K>
void unique_flags(const K* k, int64_t m, KeyFmt f, int depth, int* flags, hipStream_t s) {
  hipLaunchKernelGGL(k_unique_flags<K>, dim3(blocks_for(m, 256)), dim3(256), 0, s, k, m, f, depth, flags);
}

// Sort, unique flags, ranks and per-level counts of up to kSmallSortKeys keys in one
// workgroup (bitonic sort in LDS): the same outputs as radix sort + k_unique_flags + scan +
// k_level_hist, in one launch instead of about a dozen for the few keys an incremental step
// adds.  Writes meta_unique = number of unique keys and meta_lev[l] = unique keys of level l.
// Bitonic sort of s[0, np) in LDS (np a power of two, keys past m set to ~0), then the unique
// flags, ranks and per-level counts of the first m keys: the tail of both one-workgroup sorts.
template <typename K>
__device__ __forceinline__ void sort_unique_lds(K* s, int m, int np, KeyFmt f, int depth, K* __restrict__ out,
                                                int* __restrict__ flags, int* __restrict__ rank,
                                                int* __restrict__ meta_unique, int* __restrict__ meta_lev,
                                                int* __restrict__ meta_zero, int* lev_h, int* ws) {
  const int t = threadIdx.x;
  if (t < kMaxLev) lev_h[t] = 0;
  __syncthreads();
  for (int k = 2; k <= np; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = t; i < np; i += 1024) {
        const int p = i ^ j;
        if (p > i) {
          const K a = s[i], b = s[p];
          if ((a > b) == ((i & k) == 0)) {
            s[i] = b;
            s[p] = a;
          }
        }
      }
      __syncthreads();
    }
  // thread t owns keys [4t, 4t + 4)
  int fl[4], sum = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = 4 * t + q;
    fl[q] = 0;
    if (i < m) {
      const K v = s[i];
      const int lev = (int)(v >> (3 * f.b));
      fl[q] = ((i == 0 || s[i - 1] != v) && lev < depth) ? 1 : 0;
      if (fl[q]) atomicAdd(&lev_h[lev], 1);
      out[i] = v;
      flags[i] = fl[q];
    }
    sum += fl[q];
  }
  // exclusive scan of the per-thread sums over 16 waves
  const int lane = t & 63, wave = t >> 6;
  int x = sum;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) ws[wave] = x;
  __syncthreads();
  int pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 16; ++w) {
    pre += w < wave ? ws[w] : 0;
    tot += ws[w];
  }
  pre += x - sum;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int i = 4 * t + q;
    if (i < m) rank[i] = pre;
    pre += fl[q];
  }
  if (t == 0) {
    *meta_unique = tot;
    *meta_zero = 0;  // a counter the caller accumulates next
  }
  if (t < kMaxLev) meta_lev[t] = lev_h[t];
}

template <typename K>
__global__ void __launch_bounds__(1024) k_sort_unique_small(const K* __restrict__ in, int m, KeyFmt f, int depth,
                                                            K* __restrict__ out, int* __restrict__ flags,
                                                            int* __restrict__ rank, int* __restrict__ meta_unique,
                                                            int* __restrict__ meta_lev, int* __restrict__ meta_zero,
                                                            const int* __restrict__ mdev) {
  __shared__ K s[kSmallSortKeys];
  __shared__ int lev_h[kMaxLev];
  __shared__ int ws[16];
  const int t = threadIdx.x;
  if (mdev) m = *mdev;  // the key count a kernel before wrote (uniform over the workgroup)
  if (m <= 0 || m > kSmallSortKeys) {  // nothing sorted: the caller sees meta_unique = -1
    if (t == 0) {
      *meta_unique = -1;
      *meta_zero = 0;
    }
    return;
  }
  int np = 2;
  while (np < m) np <<= 1;
  for (int i = t; i < np; i += 1024) s[i] = i < m ? in[i] : ~(K)0;
  sort_unique_lds<K>(s, m, np, f, depth, out, flags, rank, meta_unique, meta_lev, meta_zero, lev_h, ws);
}

template <typename K>
void sort_unique_small(const K* in, int64_t m, KeyFmt f, int depth, K* out, int* flags, int* rank, int* meta_unique,
                       int* meta_lev, int* meta_zero, const int* mdev, hipStream_t s) {
  hipLaunchKernelGGL(k_sort_unique_small<K>, dim3(1), dim3(1024), 0, s, in, (int)m, f, depth, out, flags, rank,
                     meta_unique, meta_lev, meta_zero, mdev);
}

// Per-level unique counts (the levCnt atomics of gvdbFindUnique), LDS histogram per block.
template <typename K>
__global__ void __launch_bounds__(256) k_level_hist(const K* __restrict__ k, const int* __restrict__ flags, int64_t m,
                                                    KeyFmt f, int* lev_cnt) {
  __shared__ int h[kMaxLev];
  if (threadIdx.x < kMaxLev) h[threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x)
    if (flags[i]) atomicAdd(&h[(int)(k[i] >> (3 * f.b))], 1);
  __syncthreads();
  if (threadIdx.x < kMaxLev && h[threadIdx.x]) atomicAdd(&lev_cnt[threadIdx.x], h[threadIdx.x]);
}

template <typename K>
void level_hist(const K* k, const int* flags, int64_t m, KeyFmt f, int* lev_cnt, hipStream_t s) {
  hipLaunchKernelGGL(k_level_hist<K>, dim3(blocks_for(m, 256, 1024)), dim3(256), 0, s, k, flags, m, f, lev_cnt);
}

// SetupNode (V.cpp:2512-2524) for the unique keys in sorted order, appended after the
// existing nodes of each level (PoolAlloc A.cpp:163-192); inner nodes get their child list.
__device__ __forceinline__ void write_node(gvdbx_node* nd, int lev, int x, int y, int z, uint64_t childlist) {
  uint4* w = (uint4*)nd;
  uint32_t h0 = (uint32_t)lev | (1u << 8);  // lev, flags=1, prio=0, pad=0
  w[0] = make_uint4(h0, (uint32_t)x, (uint32_t)y, (uint32_t)z);
  w[1] = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u);  // value = -1, vrange.x = 0
  w[2] = make_uint4(0u, 0u, (uint32_t)kUndefL, 0u);               // vrange.yz, parent = ID_UNDEFL
  w[3] = make_uint4((uint32_t)childlist, (uint32_t)(childlist >> 32), 0u, 0u);
}

// link_leaves: every new node is a leaf whose parent is linked already (no new inner node), so
// each new leaf is also linked here (k_link_level's work for level 0, one launch fewer).
template <typename K>
__global__ void k_setup_nodes(DevTree T, const K* __restrict__ k, const int* __restrict__ flags,
                              const int* __restrict__ rank, int64_t m, KeyFmt f, AllocPlan plan, int link_leaves) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m || !flags[i]) return;
  const K v = k[i];
  const uint64_t mask = (1ull << f.b) - 1ull;
  const int lev = (int)(v >> (3 * f.b));
  const int xr = (int)((v >> (2 * f.b)) & mask), yr = (int)((v >> f.b) & mask), zr = (int)(v & mask);
  const int sh = T.lgR[lev];
  const int64_t idx = plan.base[lev] + (rank[i] - plan.levstart[lev]);
  const uint64_t cl = lev > 0 ? (1ull | ((uint64_t)lev << 8) | ((uint64_t)idx << 16)) : kUndefL;
  const int x = ((T.root[0] >> sh) + xr) << sh, y = ((T.root[1] >> sh) + yr) << sh, z = ((T.root[2] >> sh) + zr) << sh;
  gvdbx_node* nd = T.nodes[lev] + idx;
  write_node(nd, lev, x, y, z, cl);
  if (link_leaves && lev == 0) {
    const int64_t pi = descend(T, x, y, z, 1);
    if (pi < 0) return;
    nd->parent = elem0(1, (uint64_t)pi);
    T.clist[1][((uint64_t)pi << (3 * T.ld[1])) + (uint64_t)child_bit(T, 1, x, y, z)] = elem0(0, (uint64_t)idx);
  }
}

template <typename K>
void setup_nodes(const DevTree& T, const K* k, const int* flags, const int* rank, int64_t m, KeyFmt f,
                 const AllocPlan& plan, hipStream_t s, bool link_leaves) {
  hipLaunchKernelGGL(k_setup_nodes<K>, dim3(blocks_for(m, 256)), dim3(256), 0, s, T, k, flags, rank, m, f, plan,
                     link_leaves ? 1 : 0);
}

__global__ void k_setup_root(DevTree T) {
  const int d = T.depth;
  write_node(T.nodes[d], d, T.root[0], T.root[1], T.root[2], 1ull | ((uint64_t)d << 8));
}
void setup_root(const DevTree& T, hipStream_t s) { hipLaunchKernelGGL(k_setup_root, dim3(1), dim3(1), 0, s, T); }

// ---------------------------------------------------------------- full rebuild by occupancy maps
// The nodes of a full rebuild are, per level below the root, the sorted unique covering nodes of
// the valid points (ActivateBricksGPU) followed by the sorted unique extra-pass box nodes that
// are not among them (ActivateExtraBricksGPU).  Instead of sorting one key per point and level,
// each level gets two occupancy maps over a box of its root-relative node coordinates, one byte
// per node, laid out x-major like the compact keys, so map order is key order: `home` marks the
// covering nodes, `box` the box nodes (a superset: a point's box holds its own voxel).  Marks are
// plain byte stores of 1 (every writer stores the same value, and the caches write back only the
// bytes they hold dirty), so no mark needs an atomic.  Each group of 32 nodes is then read as a
// 32-bit mask; counting and ranking the set bits gives every node its pool index, in the order
// the sorts give.

__device__ __forceinline__ int64_t omap_index(const OmapLevels& M, int l, int xr, int yr, int zr) {
  const unsigned x = (unsigned)(xr - M.lo[l][0]), y = (unsigned)(yr - M.lo[l][1]), z = (unsigned)(zr - M.lo[l][2]);
  if (x >= (unsigned)M.n[l][0] || y >= (unsigned)M.n[l][1] || z >= (unsigned)M.n[l][2]) return -1;
  return M.word0[l] * 32 + ((int64_t)x * M.n[l][1] + y) * M.n[l][2] + z;
}

// Marks leaf-level node (x, y, z) (voxel corner, inside the root); the plain read skips the store
// where the mark is there already (a stale read only costs a store of the same value).  A node
// outside the map box sets the overflow flag (the host then sorts).
__device__ __forceinline__ void omap_mark_leaf(const DevTree& T, const OmapLevels& M, uint8_t* map, int x, int y, int z,
                                               int* overflow) {
  const int sh = T.lgR[0];
  const int64_t b =
      omap_index(M, 0, (x >> sh) - (T.root[0] >> sh), (y >> sh) - (T.root[1] >> sh), (z >> sh) - (T.root[2] >> sh));
  if (b < 0) {
    *overflow = 1;
    return;
  }
  if (map[b] == 0) map[b] = 1;
}

// Per valid point: its covering leaf (gvdbCalcBrickId, P.cuh:352-382: trunc(q) >> log2 R0 for
// q >= 0) into `home`; with box, every leaf-level node of its extra-pass box
// [trunc(trunc(q) - r), trunc(trunc(q) + 2r)] (the corners gvdbCalcExtraBrickId enumerates,
// P.cuh:467-512, as in k_box_keys) inside the root into `box`.  The covering and box nodes of the
// inner levels are the ancestors of these (k_omap_lift).  kTwo: the box spans at most two leaf
// nodes per axis, so its up to 8 corners are marked unrolled, their loads independent.
template <bool kTwo>
__global__ void __launch_bounds__(256) k_omap_mark(DevTree T, PointsView P, OmapLevels M, float r, int box,
                                                   uint8_t* __restrict__ home, uint8_t* __restrict__ boxmap,
                                                   int* __restrict__ overflow) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const float3 p = load_f3(P.pos, i, P.pos_off, P.pos_stride);
  const float q[3] = {p.x + P.tx, p.y + P.ty, p.z + P.tz};
  if (q[0] < 0 || q[1] < 0 || q[2] < 0) return;
  const int iq[3] = {(int)q[0], (int)q[1], (int)q[2]};
  omap_mark_leaf(T, M, home, iq[0], iq[1], iq[2], overflow);
  if (!box) return;
  const int R0 = 1 << T.lgR[0];
  int lo[3], hi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float tq = (float)iq[a];
    lo[a] = trunc_multiple((int)(tq - r), T.lgR[0]);
    hi[a] = trunc_multiple((int)(tq + 2.0f * r), T.lgR[0]);
  }
  if (kTwo) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int x = lo[0] + (k & 1) * R0, y = lo[1] + ((k >> 1) & 1) * R0, z = lo[2] + (k >> 2) * R0;
      if (x <= hi[0] && y <= hi[1] && z <= hi[2] && in_root(T, x, y, z)) omap_mark_leaf(T, M, boxmap, x, y, z, overflow);
    }
    return;
  }
  for (int x = lo[0]; x <= hi[0]; x += R0)
    for (int y = lo[1]; y <= hi[1]; y += R0)
      for (int z = lo[2]; z <= hi[2]; z += R0)
        if (in_root(T, x, y, z)) omap_mark_leaf(T, M, boxmap, x, y, z, overflow);
}

// Level l from level l - 1: a node is marked when one of its children is (the covering and box
// nodes of a level are the ancestors of the leaf-level ones).  One thread per (node, row of
// children along z), both maps; children outside level l - 1's box are not there to read.
__global__ void __launch_bounds__(256) k_omap_lift(DevTree T, OmapLevels M, int l, uint8_t* __restrict__ home,
                                                   uint8_t* __restrict__ boxmap) {
  const int ld = T.ld[l], res = 1 << ld;
  const int64_t nodes = (int64_t)M.n[l][0] * M.n[l][1] * M.n[l][2];
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nodes * res * res) return;
  const int64_t e = t / (res * res);
  const int row = (int)(t - e * res * res);
  const int nz = M.n[l][2], ny = M.n[l][1];
  const int zr = (int)(e % nz) + M.lo[l][2], yr = (int)((e / nz) % ny) + M.lo[l][1],
            xr = (int)(e / ((int64_t)nz * ny)) + M.lo[l][0];
  const int cx = (xr << ld) + row / res, cy = (yr << ld) + row % res;
  bool h = false, b = false;
  for (int k = 0; k < res; ++k) {
    const int64_t c = omap_index(M, l - 1, cx, cy, (zr << ld) + k);
    if (c < 0) continue;
    h = h || home[c] != 0;
    b = b || (boxmap && boxmap[c] != 0);
  }
  const int64_t self = M.word0[l] * 32 + e;
  if (h && home[self] == 0) home[self] = 1;
  if (b && boxmap[self] == 0) boxmap[self] = 1;
}

void omap_mark(const DevTree& T, const PointsView& P, const OmapLevels& M, float r, bool box, uint8_t* home,
               uint8_t* boxmap, int* overflow, hipStream_t s) {
  if (P.n > 0) {
    // a box [trunc(q) - r, trunc(q) + 2r] spans at most 2 leaf nodes per axis when it is shorter
    // than one leaf node
    const bool two = 3.0 * (double)r + 1.0 < (double)(1 << T.lgR[0]);
    if (two)
      hipLaunchKernelGGL(k_omap_mark<true>, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, T, P, M, r, box ? 1 : 0, home,
                         boxmap, overflow);
    else
      hipLaunchKernelGGL(k_omap_mark<false>, dim3(blocks_for(P.n, 256)), dim3(256), 0, s, T, P, M, r, box ? 1 : 0,
                         home, boxmap, overflow);
  }
  for (int l = 1; l < M.nlev; ++l) {
    const int64_t res = (int64_t)1 << T.ld[l];
    const int64_t threads = (int64_t)M.n[l][0] * M.n[l][1] * M.n[l][2] * res * res;
    hipLaunchKernelGGL(k_omap_lift, dim3(blocks_for(threads, 256)), dim3(256), 0, s, T, M, l, home, boxmap);
  }
}

// Bit j = (byte j of group w != 0), for the 32 bytes of group w.
__device__ __forceinline__ uint32_t omap_group(const uint8_t* __restrict__ map, int64_t w) {
  const uint4* g = (const uint4*)(map + w * 32);
  uint32_t m = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint4 v = g[h];
    const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if ((u[k] >> (8 * b)) & 0xFFu) m |= 1u << (h * 16 + k * 4 + b);
  }
  return m;
}

// Per group of 32 nodes: the home nodes (ch) and the box nodes that are not home nodes (cb).
__global__ void k_omap_counts(const uint8_t* __restrict__ home, const uint8_t* __restrict__ boxmap, int64_t words,
                              int* __restrict__ ch, int* __restrict__ cb) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= words) return;
  const uint32_t h = omap_group(home, w);
  ch[w] = __popc(h);
  cb[w] = boxmap ? __popc(omap_group(boxmap, w) & ~h) : 0;
}

void omap_counts(const uint8_t* home, const uint8_t* boxmap, int64_t words, int* ch, int* cb, hipStream_t s) {
  hipLaunchKernelGGL(k_omap_counts, dim3(blocks_for(words, 256)), dim3(256), 0, s, home, boxmap, words, ch, cb);
}

// Per level: the home and the other box node counts, from the exclusive scans of the group
// counts (sh, sb hold words + 1 entries: the scan with the total appended).
__global__ void k_omap_level_totals(OmapLevels M, const int* __restrict__ sh, const int* __restrict__ sb,
                                    int* __restrict__ home_cnt, int* __restrict__ box_cnt) {
  const int l = threadIdx.x;
  if (l >= M.nlev) return;
  home_cnt[l] = sh[M.word0[l + 1]] - sh[M.word0[l]];
  box_cnt[l] = sb[M.word0[l + 1]] - sb[M.word0[l]];
}

void omap_level_totals(const OmapLevels& M, const int* sh, const int* sb, int* home_cnt, int* box_cnt,
                       hipStream_t s) {
  hipLaunchKernelGGL(k_omap_level_totals, dim3(1), dim3(64), 0, s, M, sh, sb, home_cnt, box_cnt);
}

// Writes the node of every marked entry: level l's pool holds its home nodes in map order, then
// its other box nodes in map order (C.home[l] = the level's home count).  Child lists of inner
// nodes are filled by the caller, links by the link launches.
__global__ void k_omap_nodes(DevTree T, OmapLevels M, const uint8_t* __restrict__ home,
                             const uint8_t* __restrict__ boxmap, const int* __restrict__ sh,
                             const int* __restrict__ sb, OmapCounts C) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= M.word0[M.nlev]) return;
  int l = 0;
  while (w >= M.word0[l + 1]) ++l;
  const uint32_t h = omap_group(home, w), b = boxmap ? (omap_group(boxmap, w) & ~h) : 0u;
  if (!h && !b) return;
  const int sft = T.lgR[l];
  const int64_t first = M.word0[l] * 32;
  auto put = [&](int bitpos, int64_t idx) {
    const int64_t e = w * 32 + bitpos - first;  // entry inside the level's box
    const int nz = M.n[l][2], ny = M.n[l][1];
    const int zr = (int)(e % nz) + M.lo[l][2], yr = (int)((e / nz) % ny) + M.lo[l][1],
              xr = (int)(e / ((int64_t)nz * ny)) + M.lo[l][0];
    const int x = ((T.root[0] >> sft) + xr) << sft, y = ((T.root[1] >> sft) + yr) << sft,
              z = ((T.root[2] >> sft) + zr) << sft;
    const uint64_t cl = l > 0 ? (1ull | ((uint64_t)l << 8) | ((uint64_t)idx << 16)) : kUndefL;
    write_node(T.nodes[l] + idx, l, x, y, z, cl);
  };
  int64_t ih = sh[w] - sh[M.word0[l]];
  for (uint32_t m = h; m; m &= m - 1) put(__ffs(m) - 1, ih++);
  int64_t ib = C.home[l] + sb[w] - sb[M.word0[l]];
  for (uint32_t m = b; m; m &= m - 1) put(__ffs(m) - 1, ib++);
}

void omap_nodes(const DevTree& T, const OmapLevels& M, const uint8_t* home, const uint8_t* boxmap, const int* sh,
                const int* sb, const OmapCounts& C, hipStream_t s) {
  const int64_t words = M.word0[M.nlev];
  if (words <= 0) return;
  hipLaunchKernelGGL(k_omap_nodes, dim3(blocks_for(words, 256)), dim3(256), 0, s, T, M, home, boxmap, sh, sb, C);
}

// gvdbLinkBricks (N.cuh:470-505): flagged node -> parent by descent, parent's child slot.
__global__ void k_link_level(DevTree T, int lev, int64_t first, int64_t count) {
  const int64_t i = first + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= first + count) return;
  gvdbx_node* nd = T.nodes[lev] + i;
  if (!nd->flags) return;
  const int x = nd->pos[0], y = nd->pos[1], z = nd->pos[2];
  const int64_t pi = descend(T, x, y, z, lev + 1);
  if (pi < 0) return;
  nd->parent = elem0(lev + 1, (uint64_t)pi);
  T.clist[lev + 1][((uint64_t)pi << (3 * T.ld[lev + 1])) + (uint64_t)child_bit(T, lev + 1, x, y, z)] =
      elem0(lev, (uint64_t)i);
}
void link_level(const DevTree& T, int lev, int64_t first, int64_t count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_link_level, dim3(blocks_for(count, 256)), dim3(256), 0, s, T, lev, first, count);
}

__global__ void k_flags_to_markers(DevTree T, int64_t count, int* __restrict__ markers) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) markers[i] = T.nodes[0][i].flags;
}
void copy_flags_to_markers(const DevTree& T, int64_t count, int* markers, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_flags_to_markers, dim3(blocks_for(count, 256)), dim3(256), 0, s, T, count, markers);
}

// gvdbDelinkLeafBricks (N.cuh:401-429).  markers (the first nmark leaves): each leaf's flag is
// first set to its marker (the incremental pass's marker application, V.cpp:1075-1081), here
// rather than before the new nodes are allocated, since nothing in between reads old leaves' flags.
__global__ void k_delink_leaves(DevTree T, int64_t count, const int* __restrict__ markers, int64_t nmark) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  gvdbx_node* nd = T.nodes[0] + i;
  if (markers && i < nmark) nd->flags = markers[i] ? 1 : 0;
  if (nd->flags) return;
  if (nd->parent == kUndef64) return;
  const int x = nd->pos[0], y = nd->pos[1], z = nd->pos[2];
  const int64_t pi = descend(T, x, y, z, 1);
  nd->parent = kUndef64;
  if (pi < 0) return;
  T.clist[1][((uint64_t)pi << (3 * T.ld[1])) + (uint64_t)child_bit(T, 1, x, y, z)] = kUndef64;
}
void delink_leaves(const DevTree& T, int64_t count, const int* markers, int64_t nmark, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_delink_leaves, dim3(blocks_for(count, 256)), dim3(256), 0, s, T, count, markers, nmark);
}

// gvdbDelinkBricks (N.cuh:431-467): one wave per inner node scans its child list.
__global__ void __launch_bounds__(256) k_delink_level(DevTree T, int lev, int64_t count) {
  const int64_t i = item_index() * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= count) return;
  gvdbx_node* nd = T.nodes[lev] + i;
  if (nd->parent == kUndef64) {
    if (lane == 0) nd->flags = 0;
    return;
  }
  const int64_t nch = 1ll << (3 * T.ld[lev]);
  const uint64_t* cl = T.clist[lev] + ((uint64_t)i << (3 * T.ld[lev]));
  bool any = false;
  for (int64_t k = lane; k < nch; k += 64) any |= (cl[k] != kUndef64);
  any = __ballot(any) != 0;
  if (lane != 0) return;
  if (any) { nd->flags = 1; return; }
  nd->flags = 0;
  const int x = nd->pos[0], y = nd->pos[1], z = nd->pos[2];
  const int64_t pi = descend(T, x, y, z, lev + 1);
  nd->parent = kUndef64;
  if (pi < 0) return;
  T.clist[lev + 1][((uint64_t)pi << (3 * T.ld[lev + 1])) + (uint64_t)child_bit(T, lev + 1, x, y, z)] = kUndef64;
}
void delink_level(const DevTree& T, int lev, int64_t count, hipStream_t s) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_delink_level, item_grid((count + 3) / 4), dim3(256), 0, s, T, lev, count);
}

// by_parent = 0: mFlags (UpdateAtlas V.cpp:2655-2661); 1: mParent != ID_UNDEF64
// (gvdbSetFlagSubcell P.cuh:109-118).
__global__ void k_leaf_flags(DevTree T, int64_t count, int by_parent, int* __restrict__ out,
                             int* __restrict__ clear, int64_t nclear) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nclear) clear[i] = -1;
  if (i >= count) return;
  const gvdbx_node* nd = T.nodes[0] + i;
  out[i] = by_parent ? (nd->parent != kUndef64 ? 1 : 0) : (nd->flags ? 1 : 0);
}
void leaf_flags(const DevTree& T, int64_t count, int by_parent, int* out, hipStream_t s, const BrickMapOut* clear) {
  const int64_t nclear = clear ? 2 * (int64_t)clear->n[0] * clear->n[1] * clear->n[2] : 0;  // (leaf, slot) pairs
  const int64_t items = count > nclear ? count : nclear;
  if (items <= 0) return;
  hipLaunchKernelGGL(k_leaf_flags, dim3(blocks_for(items, 256)), dim3(256), 0, s, T, count, by_parent, out,
                     clear ? clear->bmap : nullptr, nclear);
}

__global__ void __launch_bounds__(256) k_count_flags(const int* __restrict__ f, int64_t n, int* out) {
  int c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += f[i] != 0;
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(out, c);
}
void count_flags(const int* flags, int64_t n, int* out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_count_flags, dim3(blocks_for(n, 256, 1024)), dim3(256), 0, s, flags, n, out);
}

// *out (zeroed by the caller) += flagged leaves among the first n; with markers, the leaves whose
// marker is set (the flags they get once the markers are applied).  Writes no node: the count is
// taken before the incremental pass is known to go ahead.
__global__ void __launch_bounds__(256) k_count_leaf_flags(DevTree T, int64_t n, const int* __restrict__ markers,
                                                         int* out) {
  int c = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    c += markers ? (markers[i] != 0) : (T.nodes[0][i].flags != 0);
  c = wave_sum(c);
  // one device atomic per workgroup (they all add to one counter)
  __shared__ int ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int b = ws[0] + ws[1] + ws[2] + ws[3];
    if (b) atomicAdd(out, b);
  }
}
void count_leaf_flags(const DevTree& T, int64_t n, const int* markers, int* out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_count_leaf_flags, dim3(blocks_for(n, 256, 1024)), dim3(256), 0, s, T, n, markers, out);
}

// ================================================================ atlas
// UpdateAtlas (V.cpp:2630-2704): slot = rank of flagged leaves; mValue = reference atlas
// corner of the slot (getAtlasPos A.cpp:705-715).  Unflagged leaves keep their stale mValue.
__device__ __forceinline__ void assign_slot(const DevTree& T, int64_t i, int s, int apron, int br, int cx, int cy,
                                            int* leaf_slot, int* slot_leaf, int4* slot_pos) {
  gvdbx_node* nd = T.nodes[0] + i;
  const int a2 = cx * cy;
  const int z = s / a2, rem = s - z * a2, y = rem / cx, x = rem - y * cx;
  nd->value[0] = x * br + apron;
  nd->value[1] = y * br + apron;
  nd->value[2] = z * br + apron;
  leaf_slot[i] = s;
  slot_leaf[s] = (int)i;
  slot_pos[s] = make_int4(nd->pos[0], nd->pos[1], nd->pos[2], (int)i);
}

__global__ void k_assign_slots(DevTree T, const int* __restrict__ flags, const int* __restrict__ rank, int64_t count,
                               int apron, int br, int cx, int cy, int* __restrict__ leaf_slot,
                               int* __restrict__ slot_leaf, int4* __restrict__ slot_pos, BrickMapOut tab) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  if (!flags[i]) { leaf_slot[i] = -1; return; }
  assign_slot(T, i, rank[i], apron, br, cx, cy, leaf_slot, slot_leaf, slot_pos);
  if (tab.bmap) {  // the leaf's cell of the lookup table
    const gvdbx_node* nd = T.nodes[0] + i;
    const unsigned bx = (unsigned)((nd->pos[0] >> tab.lg0) - tab.lo[0]),
                   by = (unsigned)((nd->pos[1] >> tab.lg0) - tab.lo[1]),
                   bz = (unsigned)((nd->pos[2] >> tab.lg0) - tab.lo[2]);
    if (bx < (unsigned)tab.n[0] && by < (unsigned)tab.n[1] && bz < (unsigned)tab.n[2])
      reinterpret_cast<int2*>(tab.bmap)[((int64_t)bz * tab.n[1] + by) * tab.n[0] + bx] = make_int2((int)i, rank[i]);
  }
}

void assign_slots(const DevTree& T, const int* flags, const int* rank, int64_t count, int apron, int br, int cx, int cy,
                  int* leaf_slot, int* slot_leaf, int4* slot_pos, hipStream_t s, const BrickMapOut* table) {
  if (count <= 0) return;
  BrickMapOut tab;
  if (table) tab = *table;
  else memset(&tab, 0, sizeof(tab));
  hipLaunchKernelGGL(k_assign_slots, dim3(blocks_for(count, 256)), dim3(256), 0, s, T, flags, rank, count, apron, br,
                     cx, cy, leaf_slot, slot_leaf, slot_pos, tab);
}

// ================================================================ subcells
// (count / insert live in binning.hip)

// gvdbCalcSubcellPos (P.cuh:677-694).
__global__ void k_sc_positions(DevTree T, int64_t leaves, SubcellCfg c, const int* __restrict__ map,
                               int* __restrict__ sc_pos, int* __restrict__ sc_nid, int* __restrict__ map_leaf,
                               int* __restrict__ ocnt, unsigned long long* __restrict__ vsum) {
  // one thread per (leaf, subcell), grid-strided
  const int per = c.D * c.D * c.D;
  const int64_t total = leaves * per;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / per;
    const int s = (int)(t - i * per);
    const gvdbx_node* nd = T.nodes[0] + i;
    if (nd->parent == kUndef64) continue;
    const int mi = map[i];
    if (s == 0) map_leaf[mi] = (int)i;
    const int64_t k = (int64_t)mi * per + s;
    sc_pos[3 * k + 0] = nd->pos[0] + (s % c.D) * c.sc;
    sc_pos[3 * k + 1] = nd->pos[1] + ((s / c.D) % c.D) * c.sc;
    sc_pos[3 * k + 2] = nd->pos[2] + ((s / c.D / c.D) % c.D) * c.sc;
    sc_nid[k] = (int)i;
    if (ocnt) ocnt[k] = 0;
    if (vsum) vsum[3 * k] = vsum[3 * k + 1] = vsum[3 * k + 2] = 0ull;
  }
}
void sc_positions(const DevTree& T, int64_t leaves, SubcellCfg c, const int* map, int* sc_pos, int* sc_nid,
                  int* map_leaf, int* ocnt, unsigned long long* vsum, hipStream_t s) {
  if (leaves <= 0) return;
  const int64_t total = leaves * c.D * c.D * c.D;
  hipLaunchKernelGGL(k_sc_positions, dim3(blocks_for(total, 256, 65536)), dim3(256), 0, s, T, leaves, c, map, sc_pos,
                     sc_nid, map_leaf, ocnt, vsum);
}

// ================================================================ gathers
// gvdbGatherDensity / gvdbGatherLevelSet (P.cuh:696-773).  One block per subcell, one
// thread per voxel; the subcell's list is staged through LDS in chunks.  The reductions
// max_p(r - |p-v|) and min_p(|p-v| - r) are taken on |p-v|^2 with one sqrt at the end
// (sqrt and the subtraction are monotonic, so the result is the same reduction).

__device__ __forceinline__ int64_t atlas_voxel(int slot, int br, int lx, int ly, int lz) {
  return (((int64_t)slot * br + lz) * br + ly) * br + lx;
}

// kQ16: the list holds ushort3 positions quantized to the box [qmin, qmin + qrange]
// (InsertPointsSubcell_FP16), decoded as gvdbGatherLevelSet_fp16 does (P.cuh:791-793)
template <bool kQ16>
__global__ void __launch_bounds__(512) k_gather_field(SubcellCfg c, int64_t nsc, const int* __restrict__ sc_pos,
                                                      const int* __restrict__ sc_nid, const int* __restrict__ cnt,
                                                      const int* __restrict__ off, const float* __restrict__ list_pos,
                                                      const uint16_t* __restrict__ list_q16, float3 qmin,
                                                      float3 qrange, const int* __restrict__ leaf_slot, int apron,
                                                      float* __restrict__ chan, int accum, int mode, float init) {
  __shared__ float sx[512], sy[512], sz[512];
  const int64_t k = item_index();
  if (k >= nsc) return;
  const int nt = c.sc * c.sc * c.sc;
  const int t = threadIdx.x;
  const int tx = t % c.sc, ty = (t / c.sc) % c.sc, tz = t / (c.sc * c.sc);
  const float vx = (float)((sc_pos[3 * k + 0] & c.cmask) + tx), vy = (float)((sc_pos[3 * k + 1] & c.cmask) + ty),
              vz = (float)((sc_pos[3 * k + 2] & c.cmask) + tz);
  const int n = cnt[k], o = off[k];
  float md2 = INFINITY;
  for (int base = 0; base < n; base += nt) {
    const int j = base + t;
    if (j < n) {
      if (kQ16) {
        sx[t] = dequant16(list_q16[3 * (int64_t)(o + j) + 0], qmin.x, qrange.x);
        sy[t] = dequant16(list_q16[3 * (int64_t)(o + j) + 1], qmin.y, qrange.y);
        sz[t] = dequant16(list_q16[3 * (int64_t)(o + j) + 2], qmin.z, qrange.z);
      } else {
        sx[t] = list_pos[3 * (int64_t)(o + j) + 0];
        sy[t] = list_pos[3 * (int64_t)(o + j) + 1];
        sz[t] = list_pos[3 * (int64_t)(o + j) + 2];
      }
    }
    __syncthreads();
    const int lim = min(nt, n - base);
    for (int jj = 0; jj < lim; ++jj) {
      const float dx = sx[jj] - vx, dy = sy[jj] - vy, dz = sz[jj] - vz;
      md2 = fminf(md2, dist2(dx, dy, dz));
    }
    __syncthreads();
  }
  const int br = c.res0 + 2 * apron;
  const int m = c.res0 - 1;
  const int slot = leaf_slot[sc_nid[k]];
  if (slot < 0) return;
  const int64_t vi = atlas_voxel(slot, br, ((int)vx & m) + apron, ((int)vy & m) + apron, ((int)vz & m) + apron);
  float val;
  if (mode == 0) {
    val = accum ? chan[vi] : 0.0f;
    if (n > 0) val = fmaxf(val, c.r - sqrtf(md2));
  } else {
    val = accum ? chan[vi] : init;
    if (n > 0) val = fminf(val, sqrtf(md2) - c.r);
  }
  chan[vi] = val;
}

void gather_field(SubcellCfg c, int64_t nsc, const int* sc_pos, const int* sc_nid, const int* cnt, const int* off,
                  const float* list_pos, const int* leaf_slot, int apron, float* chan, int accum, int mode, float init,
                  hipStream_t s) {
  if (nsc <= 0) return;
  hipLaunchKernelGGL(k_gather_field<false>, item_grid(nsc), dim3(c.sc * c.sc * c.sc), 0, s, c, nsc, sc_pos, sc_nid,
                     cnt, off, list_pos, (const uint16_t*)nullptr, make_float3(0.f, 0.f, 0.f),
                     make_float3(1.f, 1.f, 1.f), leaf_slot, apron, chan, accum, mode, init);
}

void levelset_field_q16(SubcellCfg c, int64_t nsc, const int* sc_pos, const int* sc_nid, const int* cnt,
                        const int* off, const uint16_t* list_q16, const float qmin[3], const float qrange[3],
                        const int* leaf_slot, int apron, float* chan, float init, hipStream_t s) {
  if (nsc <= 0) return;
  hipLaunchKernelGGL(k_gather_field<true>, item_grid(nsc), dim3(c.sc * c.sc * c.sc), 0, s, c, nsc, sc_pos, sc_nid,
                     cnt, off, (const float*)nullptr, list_q16, make_float3(qmin[0], qmin[1], qmin[2]),
                     make_float3(qrange[0], qrange[1], qrange[2]), leaf_slot, apron, chan, 0, 1, init);
}

// Level set over 4^3 subcells (GatherLevelSet, P.cuh:736-773, and ScatterReduceLevelSet, which is
// the same segmented min over the (point, subcell) pairs the insert sorted by subcell): one wave
// per subcell, each lane a block of 16 voxels over a sixteenth of the subcell's pair segment.
// Each voxel keeps min |p - v|^2 and takes one square root at the end: sqrtf and the
// subtraction of r are monotone, so sqrtf(min d2) - r equals the per-pair min of
// sqrtf(d2) - r bit for bit.
__global__ void __launch_bounds__(256) k_levelset_sc4(SubcellCfg c, int64_t nsc, const int* __restrict__ sc_pos,
                                                     const int* __restrict__ sc_nid, const int* __restrict__ cnt,
                                                     const int* __restrict__ off, const float* __restrict__ list_pos,
                                                     const int* __restrict__ leaf_slot, int apron,
                                                     float* __restrict__ chan, int accum, float init) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t k = item_index() * 4 + wave;
  if (k >= nsc) return;
  const int lane = threadIdx.x & 63;
  const int bx = sc_pos[3 * k] & c.cmask, by = sc_pos[3 * k + 1] & c.cmask, bz = sc_pos[3 * k + 2] & c.cmask;
  const int n = cnt[k];
  const float* __restrict__ lp = list_pos + 3 * (int64_t)off[k];
  // Lane (b, g): the 2 x 2 block b of voxel columns (x, y) of the subcell, all four z, over the
  // entries j = g (mod 16).  |p - v|^2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx)) (dist2's
  // operations, same bits): per entry the lane forms dx and dx * dx for its two x, dy for its
  // two y, the four inner fmaf(dy, dy, dx * dx) and dz for the four z, so its 16 voxels cost
  // 2.4 VALU instructions each instead of 7.  The 16 lane groups read 16 different entries per
  // load (each address shared by 4 lanes); their minima are combined at the end.
  const int b = lane & 3, g = lane >> 2;
  const int ix0 = bx + 2 * (b & 1), iy0 = by + 2 * (b >> 1);
  const float vx0 = (float)ix0, vx1 = (float)(ix0 + 1), vy0 = (float)iy0, vy1 = (float)(iy0 + 1);
  const float vz0 = (float)bz;
  float mn[16];  // voxel (x0 + xo, y0 + yo, z0 + z) at ((yo * 2 + xo) * 4 + z)
#pragma unroll
  for (int v = 0; v < 16; ++v) mn[v] = INFINITY;
  auto take = [&](float px, float py, float pz) {
    const float dx0 = px - vx0, dx1 = px - vx1, dy0 = py - vy0, dy1 = py - vy1;
    const float sx0 = dx0 * dx0, sx1 = dx1 * dx1;
    const float t[4] = {fmaf(dy0, dy0, sx0), fmaf(dy0, dy0, sx1), fmaf(dy1, dy1, sx0), fmaf(dy1, dy1, sx1)};
    float dz[4];
#pragma unroll
    for (int z = 0; z < 4; ++z) dz[z] = pz - (vz0 + (float)z);
#pragma unroll
    for (int cxy = 0; cxy < 4; ++cxy)
#pragma unroll
      for (int z = 0; z < 4; ++z) mn[cxy * 4 + z] = fminf(mn[cxy * 4 + z], fmaf(dz[z], dz[z], t[cxy]));
  };
  int jb = 0;
  constexpr int kStep = 32;  // entries per step: two per lane group, loaded one step ahead
  if (n >= kStep) {
    float nx[6];
#pragma unroll
    for (int e = 0; e < 2; ++e) {
      const float* q = lp + 3 * (16 * e + g);
      nx[3 * e] = q[0];
      nx[3 * e + 1] = q[1];
      nx[3 * e + 2] = q[2];
    }
    for (; jb + kStep <= n; jb += kStep) {
      float p[6];
#pragma unroll
      for (int e = 0; e < 6; ++e) p[e] = nx[e];
      if (jb + 2 * kStep <= n) {
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const float* q = lp + 3 * (jb + kStep + 16 * e + g);
          nx[3 * e] = q[0];
          nx[3 * e + 1] = q[1];
          nx[3 * e + 2] = q[2];
        }
      }
      take(p[0], p[1], p[2]);
      take(p[3], p[4], p[5]);
    }
  }
  for (; jb < n; jb += 16) {
    const int j = jb + g;
    if (j < n) take(lp[3 * j], lp[3 * j + 1], lp[3 * j + 2]);
  }
  // combine the 16 entry phases (lane bits 2..5)
#pragma unroll
  for (int v = 0; v < 16; ++v) {
    mn[v] = fminf(mn[v], __shfl_xor(mn[v], 4, 64));
    mn[v] = fminf(mn[v], __shfl_xor(mn[v], 8, 64));
    mn[v] = fminf(mn[v], __shfl_xor(mn[v], 16, 64));
    mn[v] = fminf(mn[v], __shfl_xor(mn[v], 32, 64));
  }
  // lane (b, g) writes voxel g of its block: xo = g & 1, yo = (g >> 1) & 1, z = g >> 2
  const int xo = g & 1, yo = (g >> 1) & 1, zo = g >> 2;
  const int sel = (yo * 2 + xo) * 4 + zo;
  float md2 = mn[0];
#pragma unroll
  for (int v = 1; v < 16; ++v) md2 = sel == v ? mn[v] : md2;
  const int slot = leaf_slot[sc_nid[k]];
  if (slot < 0) return;
  const int br = c.res0 + 2 * apron, m = c.res0 - 1;
  const int64_t vi =
      atlas_voxel(slot, br, ((ix0 + xo) & m) + apron, ((iy0 + yo) & m) + apron, ((bz + zo) & m) + apron);
  float val = accum ? chan[vi] : init;
  if (n > 0) val = fminf(val, sqrtf(md2) - c.r);
  chan[vi] = val;
}

void levelset_field(SubcellCfg c, int64_t nsc, const int* sc_pos, const int* sc_nid, const int* cnt, const int* off,
                    const float* list_pos, const int* leaf_slot, int apron, float* chan, int accum, float init,
                    hipStream_t s) {
  if (nsc <= 0) return;
  if (c.sc != 4) {
    gather_field(c, nsc, sc_pos, sc_nid, cnt, off, list_pos, leaf_slot, apron, chan, accum, 1, init, s);
    return;
  }
  hipLaunchKernelGGL(k_levelset_sc4, item_grid((nsc + 3) / 4), dim3(256), 0, s, c, nsc, sc_pos, sc_nid,
                     cnt, off, list_pos, leaf_slot, apron, chan, accum, init);
}

// ================================================================ trilinear sampling
// getVolSampleF (R.cuh:116-125): leaf at the point, tex3D-style linear filter with voxel
// v centred at v + 0.5, read from the brick's padded (apron) voxels; 0 outside leaves.
__global__ void k_sample(DevTree T, const int* __restrict__ leaf_slot, int res0, int apron,
                         const float* __restrict__ chan, const float* __restrict__ q, int64_t n,
                         float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = q[3 * i], py = q[3 * i + 1], pz = q[3 * i + 2];
  if (!(px >= 0.f && py >= 0.f && pz >= 0.f)) { out[i] = 0.f; return; }
  const int ix = (int)px, iy = (int)py, iz = (int)pz;
  const int slot = find_slot(T, ix, iy, iz, leaf_slot);
  if (slot < 0) { out[i] = 0.f; return; }
  const int m = ~(res0 - 1);
  const float xb = px - (float)(ix & m) + (float)apron - 0.5f;
  const float yb = py - (float)(iy & m) + (float)apron - 0.5f;
  const float zb = pz - (float)(iz & m) + (float)apron - 0.5f;
  const float fx = floorf(xb), fy = floorf(yb), fz = floorf(zb);
  const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
  const float ax = xb - fx, ay = yb - fy, az = zb - fz;
  const int br = res0 + 2 * apron;
  const float* b = chan + (int64_t)slot * br * br * br;
  auto at = [&](int x, int y, int z) { return b[(z * br + y) * br + x]; };
  const float c00 = at(x0, y0, z0) * (1.f - ax) + at(x0 + 1, y0, z0) * ax;
  const float c10 = at(x0, y0 + 1, z0) * (1.f - ax) + at(x0 + 1, y0 + 1, z0) * ax;
  const float c01 = at(x0, y0, z0 + 1) * (1.f - ax) + at(x0 + 1, y0, z0 + 1) * ax;
  const float c11 = at(x0, y0 + 1, z0 + 1) * (1.f - ax) + at(x0 + 1, y0 + 1, z0 + 1) * ax;
  const float c0 = c00 * (1.f - ay) + c10 * ay, c1 = c01 * (1.f - ay) + c11 * ay;
  out[i] = c0 * (1.f - az) + c1 * az;
}

void sample_trilinear(const DevTree& T, const int* leaf_slot, int res0, int apron, const float* chan, const float* q,
                      int64_t n, float* out, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sample, dim3(blocks_for(n, 256)), dim3(256), 0, s, T, leaf_slot, res0, apron, chan, q, n, out);
}

// explicit instantiations
template void brick_keys<uint32_t>(const DevTree&, const PointsView&, KeyFmt, int, uint32_t*, hipStream_t);
template void brick_keys<uint64_t>(const DevTree&, const PointsView&, KeyFmt, int, uint64_t*, hipStream_t);
template void expand_ancestors<uint32_t>(const DevTree&, const uint32_t*, const int*, const int*, int64_t, KeyFmt,
                                         uint32_t*, hipStream_t);
template void expand_ancestors<uint64_t>(const DevTree&, const uint64_t*, const int*, const int*, int64_t, KeyFmt,
                                         uint64_t*, hipStream_t);
template void box_keys<uint32_t>(const DevTree&, const PointsView&, KeyFmt, float, float, uint32_t*, int64_t, int*,
                                 int*, hipStream_t, uint32_t*, int, int*);
template void box_keys<uint64_t>(const DevTree&, const PointsView&, KeyFmt, float, float, uint64_t*, int64_t, int*,
                                 int*, hipStream_t, uint64_t*, int, int*);
template void compact_regions<uint32_t>(const uint32_t*, int64_t, const int*, const int*, int64_t, uint32_t*,
                                        int64_t, hipStream_t);
template void compact_regions<uint64_t>(const uint64_t*, int64_t, const int*, const int*, int64_t, uint64_t*,
                                        int64_t, hipStream_t);
template void sort_unique_small<uint32_t>(const uint32_t*, int64_t, KeyFmt, int, uint32_t*, int*, int*, int*, int*,
                                          int*, const int*, hipStream_t);
template void sort_unique_small<uint64_t>(const uint64_t*, int64_t, KeyFmt, int, uint64_t*, int*, int*, int*, int*,
                                          int*, const int*, hipStream_t);
template void unique_flags<uint32_t>(const uint32_t*, int64_t, KeyFmt, int, int*, hipStream_t);
template void unique_flags<uint64_t>(const uint64_t*, int64_t, KeyFmt, int, int*, hipStream_t);
template void level_hist<uint32_t>(const uint32_t*, const int*, int64_t, KeyFmt, int*, hipStream_t);
template void level_hist<uint64_t>(const uint64_t*, const int*, int64_t, KeyFmt, int*, hipStream_t);
template void setup_nodes<uint32_t>(const DevTree&, const uint32_t*, const int*, const int*, int64_t, KeyFmt,
                                    const AllocPlan&, hipStream_t, bool);
template void setup_nodes<uint64_t>(const DevTree&, const uint64_t*, const int*, const int*, int64_t, KeyFmt,
                                    const AllocPlan&, hipStream_t, bool);

// ------------------------------------------------------------------ displacement

// q += v * dt in place: an FP32 product, then an FP32 sum (two roundings, no fused multiply-add),
// so the positions equal those of the same arithmetic in FP32 on the host.
__global__ void __launch_bounds__(256) k_advance(char* __restrict__ pos, int pos_off, int pos_stride,
                                                const char* __restrict__ vel, int vel_off, int vel_stride,
                                                int64_t n, float dt) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    float* q = (float*)(pos + i * pos_stride + pos_off);
    const float* v = (const float*)(vel + i * vel_stride + vel_off);
#pragma unroll
    for (int a = 0; a < 3; ++a) q[a] = __fadd_rn(q[a], __fmul_rn(v[a], dt));
  }
}

// Packed float3 arrays, both 16-byte aligned: the same arithmetic over m floats, four per lane.
__global__ void __launch_bounds__(256) k_advance_flat(float* __restrict__ q, const float* __restrict__ v, int64_t m,
                                                     float dt) {
  const int64_t m4 = m / 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m4; i += stride) {
    float4 a = reinterpret_cast<float4*>(q)[i];
    const float4 b = reinterpret_cast<const float4*>(v)[i];
    a.x = __fadd_rn(a.x, __fmul_rn(b.x, dt));
    a.y = __fadd_rn(a.y, __fmul_rn(b.y, dt));
    a.z = __fadd_rn(a.z, __fmul_rn(b.z, dt));
    a.w = __fadd_rn(a.w, __fmul_rn(b.w, dt));
    reinterpret_cast<float4*>(q)[i] = a;
  }
  const int64_t i = 4 * m4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // the last m % 4 floats
  if (i < m) q[i] = __fadd_rn(q[i], __fmul_rn(v[i], dt));
}

__global__ void __launch_bounds__(64) k_meta_export(const int* __restrict__ meta, int* host, int n, unsigned seq) {
  const int t = threadIdx.x;
  if (t < n) host[t] = meta[t];
  __threadfence_system();  // each lane's counter store completes before the barrier
  __syncthreads();
  if (t == 0) __hip_atomic_store(reinterpret_cast<unsigned*>(host + n), seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void meta_export(const int* meta, int* host, int n, unsigned seq, hipStream_t s) {
  hipLaunchKernelGGL(k_meta_export, dim3(1), dim3(64), 0, s, meta, host, n, seq);
}

void advance_points(char* pos, int pos_off, int pos_stride, const char* vel, int vel_off, int vel_stride, int64_t n,
                    float dt, hipStream_t s) {
  if (n <= 0) return;
  char* q = pos + pos_off;
  const char* v = vel + vel_off;
  if (pos_stride == 12 && vel_stride == 12 && ((uintptr_t)q & 15) == 0 && ((uintptr_t)v & 15) == 0) {
    const int64_t m = 3 * n;
    int64_t g = (m / 4 + 255) / 256;
    if (g > 16384) g = 16384;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(k_advance_flat, dim3((unsigned)g), dim3(256), 0, s, (float*)q, (const float*)v, m, dt);
    return;
  }
  int64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_advance, dim3((unsigned)g), dim3(256), 0, s, pos, pos_off, pos_stride, vel, vel_off,
                     vel_stride, n, dt);
}

}  // namespace gvdbx
