// binning.hip -- subcell binning by home brick, and the gathers that read its results.
//
// Semantics restated from gvdbCountSubcell / gvdbInsertSubcell (P.cuh:218-315): a point q
// enters every subcell whose corner lies in [trunc(q-r)/sc*sc, trunc(q+1+r)/sc*sc] and whose
// leaf exists.  The reference issues one global atomic per (point, subcell) pair, in point
// order.  Here the points are first sorted by home brick (LSD radix sort of the home key),
// then one workgroup per brick counts / inserts the pairs of its own points into LDS counters
// covering the brick's subcells plus a one-subcell shell, and flushes them with one global
// atomic per touched subcell (at most 8 workgroups touch a subcell).  Velocity sums for the
// subcell means are exact 64-bit fixed point, so their value is independent of order.
#include <limits.h>

#include "binning.hpp"
#include "primitives.hpp"

namespace gvdbx {

static inline int nblocks(int64_t n, int bs) {
  int64_t g = (n + bs - 1) / bs;
  return (int)(g < 1 ? 1 : (g > 0x7FFFFFFF ? 0x7FFFFFFF : g));
}

constexpr float kVelScale = 16777216.0f;  // 2^24 fixed-point units
__device__ __forceinline__ long long vel_fixed(float v) { return __float2ll_rn(v * kVelScale); }

// Subcell corner range of a point (the exact FP32 operation order of P.cuh:238-243); sc = 2^lgsc.
__device__ __forceinline__ void subcell_range(const float q[3], float r, int lgsc, int lo[3], int hi[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float lq = q[a] - r;
    float hq = q[a] + 1.0f;
    hq = hq + r;
    lo[a] = trunc_multiple((int)lq, lgsc);
    hi[a] = trunc_multiple((int)hq, lgsc);
  }
}

// ------------------------------------------------------------------ bounding box
// Also clears clear0[0, n0) and clear1[0, n1) (the step's counters and markers, one launch fewer).
__global__ void __launch_bounds__(256) k_bbox_partial(PointsView P, float* __restrict__ part, uint32_t* clear0,
                                                     int64_t n0, uint32_t* clear1, int64_t n1) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x, gstride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = gid; i < n0; i += gstride) clear0[i] = 0u;
  for (int64_t i = gid; i < n1; i += gstride) clear1[i] = 0u;
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < P.n; i += (int64_t)gridDim.x * blockDim.x) {
    const float3 p = load_f3(P.pos, i, P.pos_off, P.pos_stride);
    const float q[3] = {p.x + P.tx, p.y + P.ty, p.z + P.tz};
#pragma unroll
    for (int a = 0; a < 3; ++a) { mn[a] = fminf(mn[a], q[a]); mx[a] = fmaxf(mx[a], q[a]); }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], o, 64));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o, 64));
    }
  __shared__ float sm[4][6];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) { sm[w][a] = mn[a]; sm[w][3 + a] = mx[a]; }
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    float v = sm[0][a];
    for (int k = 1; k < 4; ++k) v = a < 3 ? fminf(v, sm[k][a]) : fmaxf(v, sm[k][a]);
    part[blockIdx.x * 6 + a] = v;
  }
}

// nparts <= kBBoxParts: each of the 256 threads loads its up to 4 partials at once (one memory
// round trip), then the waves reduce by shuffles and LDS
__global__ void __launch_bounds__(256) k_bbox_final(const float* __restrict__ part, int nparts, int* meta) {
  const int t = threadIdx.x;
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
  for (int k = 0; k < kBBoxParts / 256; ++k) {
    const int b = t + 256 * k;
    if (b < nparts)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        mn[a] = fminf(mn[a], part[b * 6 + a]);
        mx[a] = fmaxf(mx[a], part[b * 6 + 3 + a]);
      }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], o, 64));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], o, 64));
    }
  __shared__ float sm[4][6];
  const int w = t >> 6;
  if ((t & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      sm[w][a] = mn[a];
      sm[w][3 + a] = mx[a];
    }
  __syncthreads();
  if (t < 6) {
    float v = sm[0][t];
    for (int k = 1; k < 4; ++k) v = t < 3 ? fminf(v, sm[k][t]) : fmaxf(v, sm[k][t]);
    meta[t] = float_to_ordered(v);
  }
}

void bbox_two_pass(const PointsView& P, float* part, int* meta, hipStream_t s, uint32_t* clear0, int64_t n0,
                   uint32_t* clear1, int64_t n1) {
  const int nb = nblocks(P.n, 256) < kBBoxParts ? nblocks(P.n, 256) : kBBoxParts;
  hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, s, P, part, clear0, n0, clear1, n1);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, s, part, nb, meta);
}

// ------------------------------------------------------------------ home keys
// Key = subcell-map index of the leaf holding trunc(q); valid points whose own leaf is missing
// get the sentinel nmap and are binned by the per-point path; invalid and NOHIT points, which
// enter no subcell, get nmap + 1 and sort last, outside every range a pass walks.
// The home keys (no values: the sort generates the positions), and per tile of kRadixTile points
// the counts of the first radix digit (the low `nbits` bits), digit-major into
// counts[digit * ntiles + tile] for all 256 digits, as k_radix_up would count them.
// slot_keys: map is leaf_slot (the slot map), so the key is the point's slot, one table load
__global__ void __launch_bounds__(kKeyBlock) k_home_keys_counted(DevTree T, PointsView P, const int* __restrict__ map,
                                                          int slot_keys, int nmap, uint32_t* __restrict__ keys,
                                                          int nbits, int* __restrict__ counts, int ntiles) {
  __shared__ int hist[256];
  if (threadIdx.x < 256) hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  const uint32_t mask = (1u << nbits) - 1u;
#pragma unroll
  for (int r = 0; r < kRadixTile / kKeyBlock; ++r) {
    const int64_t i = base + r * kKeyBlock + threadIdx.x;
    if (i >= P.n) continue;
    const float3 p = load_f3(P.pos, i, P.pos_off, P.pos_stride);
    uint32_t key = (uint32_t)nmap + 1u;
    if (p.z != kNoHit) {
      const float q[3] = {p.x + P.tx, p.y + P.ty, p.z + P.tz};
      if (q[0] >= 0 && q[1] >= 0 && q[2] >= 0) {
        const int3 o = cloud_origin(P, i);
        if (slot_keys) {
          const int sl = find_slot(T, (int)q[0] + o.x, (int)q[1] + o.y, (int)q[2] + o.z, map);
          key = sl >= 0 ? (uint32_t)sl : (uint32_t)nmap;
        } else {
          const int64_t lf = find_leaf(T, (int)q[0] + o.x, (int)q[1] + o.y, (int)q[2] + o.z);
          key = lf >= 0 ? (uint32_t)map[lf] : (uint32_t)nmap;
        }
      }
    }
    keys[i] = key;
    atomicAdd(&hist[key & mask], 1);
  }
  __syncthreads();
  if (threadIdx.x < 256) counts[(int64_t)threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

void home_keys_counted(const DevTree& T, const PointsView& P, const int* map, bool slot_keys, int nmap,
                       uint32_t* keys, int bits, int* counts, hipStream_t s) {
  const int ntiles = (int)((P.n + kRadixTile - 1) / kRadixTile);
  hipLaunchKernelGGL(k_home_keys_counted, dim3(ntiles), dim3(kKeyBlock), 0, s, T, P, map, slot_keys ? 1 : 0, nmap,
                     keys, radix_first_digit_bits(bits), counts, ntiles);
}

// Batched clouds: cloud index of every point (off = exclusive prefix of the cloud sizes) and
// a flag when a valid point (all coordinates >= 0) reaches `limit` in some coordinate, i.e.
// leaves the part of its cell that keeps clouds apart.
__global__ void __launch_bounds__(256) k_batch_ids(PointsView P, const int64_t* __restrict__ off, int ncloud,
                                                   float limit, uint32_t* __restrict__ cloud, int* __restrict__ flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  int lo = 0, hi = ncloud - 1;  // last cloud with off[k] <= i
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  cloud[i] = (uint32_t)lo;
  const float3 p = load_f3(P.pos, i, P.pos_off, P.pos_stride);
  if (p.x >= 0.f && p.y >= 0.f && p.z >= 0.f && (p.x >= limit || p.y >= limit || p.z >= limit)) *flag = 1;
}

void batch_ids(const PointsView& P, const int64_t* off, int ncloud, float limit, uint32_t* cloud, int* flag,
               hipStream_t s) {
  if (P.n <= 0) return;
  hipLaunchKernelGGL(k_batch_ids, dim3(nblocks(P.n, 256)), dim3(256), 0, s, P, off, ncloud, limit, cloud, flag);
}

// start[k] = first sorted position with key >= k, for k in [0, nmap + 1]: one lower-bound
// search per key, so the work per thread is log2(n) whatever the gaps between keys.
__global__ void k_ranges(const uint32_t* __restrict__ skeys, int64_t n, int nmap, int* __restrict__ start) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > (int64_t)nmap + 1) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)skeys[mid] < k) lo = mid + 1;
    else hi = mid;
  }
  start[k] = (int)lo;
}

void key_ranges(const uint32_t* skeys, int64_t n, int nmap, int* start, hipStream_t s) {
  hipLaunchKernelGGL(k_ranges, dim3(nblocks((int64_t)nmap + 2, 256)), dim3(256), 0, s, skeys, n, nmap, start);
}

// The same starts from the boundaries of the sorted keys (keys in [0, nmax]): sorted position j
// is the start of every key in (skeys[j - 1], skeys[j]] (j = 0: [0, skeys[0]]; j = n: up to
// nmax + 1), so each start is written once, by one pass over the keys -- cheaper than a search
// per key when there are about as many keys as points.  A lane fills a short run of keys
// itself; runs longer than a wave (a block of empty keys) are filled by the whole wave.
__global__ void __launch_bounds__(256) k_key_starts(const uint32_t* __restrict__ skeys, int64_t n, int nmax,
                                                    int* __restrict__ start) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = j <= n;
  int64_t a = 0, b = -1;
  if (live) {
    a = j == 0 ? 0 : (int64_t)skeys[j - 1] + 1;
    b = j == n ? (int64_t)nmax + 1 : (int64_t)skeys[j];
  }
  const bool wide = b - a >= 64;
  if (!wide)
    for (int64_t k = a; k <= b; ++k) start[k] = (int)j;
  uint64_t pending = __ballot(wide);
  while (pending) {
    const int l = __ffsll((unsigned long long)pending) - 1;
    const int64_t la = __shfl(a, l, 64), lb = __shfl(b, l, 64), lj = __shfl(j, l, 64);
    for (int64_t k = la + lane_id(); k <= lb; k += 64) start[k] = (int)lj;
    pending &= pending - 1;
  }
}

void key_starts(const uint32_t* skeys, int64_t n, int nmax, int* start, hipStream_t s) {
  hipLaunchKernelGGL(k_key_starts, dim3(nblocks(n + 1, 256)), dim3(256), 0, s, skeys, n, nmax, start);
}

// ------------------------------------------------------------------ per-brick binning
// Window: the brick's D^3 subcells plus one subcell on every side, W = D + 2 per axis; a point
// of the brick reaches outside the window only when r + 1 > sc (those pairs go to global
// memory by tree descent).  Counting: when D <= 2 every point adds itself once to a
// (home subcell, box class) bin -- per axis its subcell range is [home + dl, home + dh] with
// dl in {-1, 0}, dh in {0, 1}, 64 classes -- and each window cell then sums the bins whose box
// covers it; otherwise pairs are counted one by one in LDS.  The brick's per-cell totals
// reserve a contiguous range of each subcell list (wgb), so the insert needs no global atomics.

struct BrickCtx {
  int bx, by, bz, W, D, D3;  // brick corner in its cloud's frame
  int ox, oy, oz;            // the cloud's cell origin
  const int* nb;
  // window cell of a corner, or -1 outside the window
  __device__ __forceinline__ int window_of(int x, int y, int z, int lgsc) const {
    const int sx = ((x - bx) >> lgsc) + 1, sy = ((y - by) >> lgsc) + 1, sz = ((z - bz) >> lgsc) + 1;
    if ((unsigned)sx >= (unsigned)W || (unsigned)sy >= (unsigned)W || (unsigned)sz >= (unsigned)W) return -1;
    return (sz * W + sy) * W + sx;
  }
  // global subcell of a window cell, or -1 when its brick does not exist
  __device__ __forceinline__ int64_t global_of(int w) const {
    const int sx = w % W - 1, sy = (w / W) % W - 1, sz = w / (W * W) - 1;
    const int ox = sx < 0 ? -1 : (sx >= D ? 1 : 0), oy = sy < 0 ? -1 : (sy >= D ? 1 : 0),
              oz = sz < 0 ? -1 : (sz >= D ? 1 : 0);
    const int b = nb[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)];
    if (b < 0) return -1;
    return (int64_t)b * D3 + ((sz - oz * D) * D + (sy - oy * D)) * D + (sx - ox * D);
  }
};

// subcell of cloud-frame corner (x, y, z) of the cloud at origin (ox, oy, oz), or -1
__device__ __forceinline__ int64_t direct_subcell(const BinArgs& A, int x, int y, int z, int ox, int oy, int oz) {
  x += ox;
  y += oy;
  z += oz;
  const int64_t lf = find_leaf(A.T, x, y, z);
  if (lf < 0) return -1;
  const int msk = A.c.res0 - 1, D = A.c.D;
  return (int64_t)A.map[lf] * (D * D * D) + ((((z & msk) >> A.c.lgsc) * D + ((y & msk) >> A.c.lgsc)) * D) +
         ((x & msk) >> A.c.lgsc);
}

__device__ __forceinline__ void add_vel_global(unsigned long long* vsum, int64_t g, const long long fv[3]) {
  atomicAdd(&vsum[3 * g + 0], (unsigned long long)fv[0]);
  atomicAdd(&vsum[3 * g + 1], (unsigned long long)fv[1]);
  atomicAdd(&vsum[3 * g + 2], (unsigned long long)fv[2]);
}

// Loads the 27 neighbour subcell-map indices (by tree descent, or from nbm where the count
// kernel of the same insert stored them); returns false for an unused map entry.
__device__ __forceinline__ bool brick_setup(const BinArgs& A, int m, int* nb, BrickCtx& B, int D, int W,
                                            bool stored = false) {
  int gx, gy, gz;
  if (A.slot_pos) {  // slot map: brick m is slot m, which always holds a leaf
    const int4 p = A.slot_pos[m];
    gx = p.x;
    gy = p.y;
    gz = p.z;
  } else {
    const int leaf = A.map_leaf[m];
    if (leaf < 0) return false;
    const gvdbx_node* nd = A.T.nodes[0] + leaf;
    gx = nd->pos[0];
    gy = nd->pos[1];
    gz = nd->pos[2];
  }
  B.bx = gx & A.c.cmask;
  B.by = gy & A.c.cmask;
  B.bz = gz & A.c.cmask;
  B.ox = gx - B.bx;
  B.oy = gy - B.by;
  B.oz = gz - B.bz;
  B.W = W;
  B.D = D;
  B.D3 = D * D * D;
  B.nb = nb;
  if (threadIdx.x < 27 && stored) {
    nb[threadIdx.x] = A.nbm[(int64_t)m * 27 + threadIdx.x];
  } else if (threadIdx.x < 27 && A.nbr) {
    // slot map: m is the leaf's slot, and the atlas neighbour table holds the slots (= map
    // indices) of the 27 bricks around it, in this order
    nb[threadIdx.x] = A.nbr[(int64_t)m * 27 + threadIdx.x];
  } else if (threadIdx.x < 27) {
    const int t = threadIdx.x;
    const int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    const int64_t lf = find_leaf(A.T, gx + dx * A.c.res0, gy + dy * A.c.res0, gz + dz * A.c.res0);
    nb[t] = lf < 0 ? -1 : A.map[lf];
  }
  return true;
}

template <int kD>
__device__ __forceinline__ void bin_count_brick(const BinArgs& A, int m, int use_bins, unsigned long long* smem64,
                                                int* nb) {
  BrickCtx B;
  const int beg = A.start[m], end = A.start[m + 1];  // in flight with the brick's set-up loads
  const int D = kD ? kD : A.c.D, W = D + 2, W3 = W * W * W, D3 = D * D * D;
  if (!brick_setup(A, m, nb, B, D, W)) {  // no leaf: nothing reaches the cell walk through it
    if (threadIdx.x < 27) A.nbm[(int64_t)m * 27 + threadIdx.x] = -1;
    return;
  }
  const int NB = use_bins ? D3 * 64 : 0;
  const bool vel = A.vsum != nullptr && A.P.vel != nullptr;
  // layout: [cvs 3*W3 u64][bvs 3*NB u64][ccnt W3 int][bcnt NB int]
  // without velocities only the counts are allocated (bin_count_smem)
  unsigned long long* cvs = smem64;
  unsigned long long* bvs = smem64 + 3 * W3;
  int* ccnt = (int*)(vel ? smem64 + 3 * W3 + 3 * NB : smem64);
  int* bcnt = ccnt + W3;
  for (int t = threadIdx.x; t < W3; t += blockDim.x) {
    ccnt[t] = 0;
    if (vel) cvs[3 * t] = cvs[3 * t + 1] = cvs[3 * t + 2] = 0ull;
  }
  for (int t = threadIdx.x; t < NB; t += blockDim.x) {
    bcnt[t] = 0;
    if (vel) bvs[3 * t] = bvs[3 * t + 1] = bvs[3 * t + 2] = 0ull;
  }
  __syncthreads();
  const SubcellCfg c = A.c;
  for (int j = beg + threadIdx.x; j < end; j += blockDim.x) {
    float q[3];
    long long fv[3] = {0, 0, 0};
    if (A.spos_out) {  // gather from the caller's arrays, and write the sorted copies
      const uint32_t i = A.svals[j];
      const float3 p = load_f3(A.P.pos, i, A.P.pos_off, A.P.pos_stride);
      q[0] = p.x + A.P.tx;
      q[1] = p.y + A.P.ty;
      q[2] = p.z + A.P.tz;
      A.spos_out[j] = make_float4(q[0], q[1], q[2], __uint_as_float(i));
      if (vel || A.svel_out) {
        const float3 v = load_f3(A.P.vel, i, A.P.vel_off, A.P.vel_stride);
        if (A.svel_out) A.svel_out[j] = make_float4(v.x, v.y, v.z, 0.f);
        fv[0] = vel_fixed(v.x);
        fv[1] = vel_fixed(v.y);
        fv[2] = vel_fixed(v.z);
      }
    } else {  // the sorted copies k_gather_sorted wrote
      const float4 p = A.spos[j];
      q[0] = p.x;
      q[1] = p.y;
      q[2] = p.z;
      if (vel) {
        const float4 v = A.svel[j];
        fv[0] = vel_fixed(v.x);
        fv[1] = vel_fixed(v.y);
        fv[2] = vel_fixed(v.z);
      }
    }
    int lo[3], hi[3];
    subcell_range(q, c.r, c.lgsc, lo, hi);
    if (use_bins) {
      const int hx = ((int)q[0] - B.bx) >> c.lgsc, hy = ((int)q[1] - B.by) >> c.lgsc, hz = ((int)q[2] - B.bz) >> c.lgsc;
      const int dlx = ((lo[0] - B.bx) >> c.lgsc) - hx, dhx = ((hi[0] - B.bx) >> c.lgsc) - hx;
      const int dly = ((lo[1] - B.by) >> c.lgsc) - hy, dhy = ((hi[1] - B.by) >> c.lgsc) - hy;
      const int dlz = ((lo[2] - B.bz) >> c.lgsc) - hz, dhz = ((hi[2] - B.bz) >> c.lgsc) - hz;
      if (dlx >= -1 && dlx <= 0 && dhx >= 0 && dhx <= 1 && dly >= -1 && dly <= 0 && dhy >= 0 && dhy <= 1 &&
          dlz >= -1 && dlz <= 0 && dhz >= 0 && dhz <= 1) {
        const int cls = (((dlz + 1) + 2 * dhz) * 4 + ((dly + 1) + 2 * dhy)) * 4 + ((dlx + 1) + 2 * dhx);
        const int bin = ((hz * D + hy) * D + hx) * 64 + cls;
        atomicAdd(&bcnt[bin], 1);
        if (vel) {
          atomicAdd(&bvs[3 * bin + 0], (unsigned long long)fv[0]);
          atomicAdd(&bvs[3 * bin + 1], (unsigned long long)fv[1]);
          atomicAdd(&bvs[3 * bin + 2], (unsigned long long)fv[2]);
        }
        continue;
      }
    }
    for (int x = lo[0]; x <= hi[0]; x += c.sc)
      for (int y = lo[1]; y <= hi[1]; y += c.sc)
        for (int z = lo[2]; z <= hi[2]; z += c.sc) {
          const int w = B.window_of(x, y, z, c.lgsc);
          if (w >= 0) {
            atomicAdd(&ccnt[w], 1);
            if (vel) {
              atomicAdd(&cvs[3 * w + 0], (unsigned long long)fv[0]);
              atomicAdd(&cvs[3 * w + 1], (unsigned long long)fv[1]);
              atomicAdd(&cvs[3 * w + 2], (unsigned long long)fv[2]);
            }
          } else {
            const int64_t g = direct_subcell(A, x, y, z, B.ox, B.oy, B.oz);
            if (g < 0) continue;
            atomicAdd(&A.ocnt[g], 1);
            if (vel) add_vel_global(A.vsum, g, fv);
          }
        }
  }
  __syncthreads();
  // scatter every populated (home, class) bin into the window cells its box covers
  for (int bin = threadIdx.x; bin < NB; bin += blockDim.x) {
    const int bc = bcnt[bin];
    if (!bc) continue;
    const int h = bin >> 6, cls = bin & 63;
    const int hx = h % D, hy = (h / D) % D, hz = h / (D * D);
    const int cx = cls & 3, cy = (cls >> 2) & 3, cz = cls >> 4;
    // class c = (dl + 1) + 2*dh  ->  range [h + dl, h + dh]
    const int x0 = hx + (cx & 1) - 1, x1 = hx + (cx >> 1);
    const int y0 = hy + (cy & 1) - 1, y1 = hy + (cy >> 1);
    const int z0 = hz + (cz & 1) - 1, z1 = hz + (cz >> 1);
    const unsigned long long v0 = vel ? bvs[3 * bin] : 0ull, v1 = vel ? bvs[3 * bin + 1] : 0ull,
                             v2 = vel ? bvs[3 * bin + 2] : 0ull;
    for (int z = z0; z <= z1; ++z)
      for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
          const int w = ((z + 1) * W + (y + 1)) * W + (x + 1);
          atomicAdd(&ccnt[w], bc);
          if (vel) {
            atomicAdd(&cvs[3 * w + 0], v0);
            atomicAdd(&cvs[3 * w + 1], v1);
            atomicAdd(&cvs[3 * w + 2], v2);
          }
        }
  }
  __syncthreads();
  if (threadIdx.x < 27) A.nbm[(int64_t)m * 27 + threadIdx.x] = nb[threadIdx.x];
  for (int w = threadIdx.x; w < W3; w += blockDim.x) {
    const int n = ccnt[w];
    int keep = 0;
    if (n > 0) {
      const int64_t g = B.global_of(w);
      if (g >= 0) {
        keep = n;
        if (vel) {  // pulled into vsum by the cell walk (no device atomics)
          unsigned long long* o = A.wvs + ((int64_t)m * W3 + w) * 3;
          o[0] = cvs[3 * w];
          o[1] = cvs[3 * w + 1];
          o[2] = cvs[3 * w + 2];
        }
      }
    }
    A.wgb[(int64_t)m * W3 + w] = keep;  // this brick's count; cell_bases turns it into a list offset
  }
}

// One list entry: the position (float3, or ushort3 quantized for the FP16 insert), the velocity
// when velocity lists are kept, the colour when colours are bound.
__device__ __forceinline__ void put_entry(const BinArgs& A, int64_t k, const float q[3], float vx, float vy, float vz,
                                          uint32_t clr) {
  if (A.list_q16) {
    A.list_q16[3 * k + 0] = quant16(q[0], A.qmin[0], A.qrange[0]);
    A.list_q16[3 * k + 1] = quant16(q[1], A.qmin[1], A.qrange[1]);
    A.list_q16[3 * k + 2] = quant16(q[2], A.qmin[2], A.qrange[2]);
  } else {
    A.list_pos[3 * k + 0] = q[0];
    A.list_pos[3 * k + 1] = q[1];
    A.list_pos[3 * k + 2] = q[2];
  }
  if (A.list_vq16) {
    A.list_vq16[3 * k + 0] = quant16(vx, A.vqmin[0], A.vqrange[0]);
    A.list_vq16[3 * k + 1] = quant16(vy, A.vqmin[1], A.vqrange[1]);
    A.list_vq16[3 * k + 2] = quant16(vz, A.vqmin[2], A.vqrange[2]);
  } else if (A.list_vel) {
    A.list_vel[3 * k + 0] = vx;
    A.list_vel[3 * k + 1] = vy;
    A.list_vel[3 * k + 2] = vz;
  }
  if (A.list_clr) A.list_clr[k] = clr;
}

template <bool kInsert>
__device__ __forceinline__ void bin_points_body(const BinArgs& A, const int* __restrict__ begin, int64_t blk,
                                                int64_t nblk) {
  const int64_t b0 = begin[0], b1 = begin[1];
  const SubcellCfg c = A.c;
  const bool vel = A.vsum != nullptr && A.P.vel != nullptr;
  for (int64_t j = b0 + blk * blockDim.x + threadIdx.x; j < b1; j += nblk * blockDim.x) {
    const int64_t i = (int64_t)A.svals[j];
    const float3 p = load_f3(A.P.pos, i, A.P.pos_off, A.P.pos_stride);
    if (p.z == kNoHit) continue;
    const float q[3] = {p.x + A.P.tx, p.y + A.P.ty, p.z + A.P.tz};
    if (q[0] < 0 || q[1] < 0 || q[2] < 0) continue;
    const int3 o = cloud_origin(A.P, i);
    float3 v = make_float3(0.f, 0.f, 0.f);
    long long fv[3] = {0, 0, 0};
    if (A.P.vel && (vel || A.list_vel || A.list_vq16)) {
      v = load_f3(A.P.vel, i, A.P.vel_off, A.P.vel_stride);
      fv[0] = vel_fixed(v.x);
      fv[1] = vel_fixed(v.y);
      fv[2] = vel_fixed(v.z);
    }
    const uint32_t clr = (kInsert && A.list_clr) ? load_u32(A.P.clr, i, A.P.clr_off, A.P.clr_stride) : 0u;
    int lo[3], hi[3];
    subcell_range(q, c.r, c.lgsc, lo, hi);
    for (int x = lo[0]; x <= hi[0]; x += c.sc)
      for (int y = lo[1]; y <= hi[1]; y += c.sc)
        for (int z = lo[2]; z <= hi[2]; z += c.sc) {
          const int64_t g = direct_subcell(A, x, y, z, o.x, o.y, o.z);
          if (g < 0) continue;
          if (!kInsert) {
            atomicAdd(&A.ocnt[g], 1);
            if (vel) add_vel_global(A.vsum, g, fv);
            continue;
          }
          const int64_t k = (int64_t)A.off[g] + A.cnt_in[g] + atomicAdd(&A.ocur[g], 1);
          put_entry(A, k, q, v.x, v.y, v.z, clr);
        }
  }
}


// The count and insert launches hold kOrphanItems more workgroups than bricks: they stride over
// the sorted points whose own leaf is missing ([start[nmap], start[nmap + 1]), rare: an insert of
// points the rebuild did not see), placing their pairs one by one with global atomics.  The
// window slices of the bricks and these one-by-one slots are disjoint, so both run in one launch.
constexpr int kOrphanItems = 256;

template <int kD, bool kPersist>
__global__ void __launch_bounds__(256) k_bin_count(BinArgs A, int nmap, int use_bins) {
  extern __shared__ unsigned long long smem64[];
  __shared__ int nb[27];
  // when the count pass gathers its bricks' points from the caller's arrays, neighbouring bricks
  // (whose points share cache lines there) run on one XCD
  for_each_brick<kPersist>(
      (int64_t)nmap + kOrphanItems, A.L.queue,
      [&](int m) {
        if (m < nmap) bin_count_brick<kD>(A, m, use_bins, smem64, nb);
        else bin_points_body<false>(A, A.start + nmap, m - nmap, kOrphanItems);
      },
      A.spos_out != nullptr);
}


// kCorners = 2 (2r + 2 <= sc): a point's subcell box [trunc(q - r), trunc(q + 1 + r)] spans under
// one subcell per axis, so it reaches at most 2 subcells per axis, all inside its brick's window;
// the up to 8 entries then take their list slots with independent LDS atomics issued back to back.
// kCorners = 3 (r + 1 <= sc, as the level set's r = 2 with subcells of 4): up to 3 subcells per
// axis, the 27 candidates likewise (a point reaching more, or a subcell outside the window, takes
// the general path's steps); 0: the general loop.
template <int kD, int kCorners>
__device__ __forceinline__ void bin_insert_brick(const BinArgs& A, int m, unsigned long long* smem64, int* nb) {
  BrickCtx B;
  const int beg = A.start[m], end = A.start[m + 1];  // in flight with the brick's set-up loads
  const int D = kD ? kD : A.c.D, W = D + 2, W3 = W * W * W;
  if (!brick_setup(A, m, nb, B, D, W, true)) return;
  int* lbase = (int*)smem64;
  int* lcnt = lbase + W3;
  __syncthreads();
  for (int w = threadIdx.x; w < W3; w += blockDim.x) {
    lbase[w] = B.global_of(w) >= 0 ? A.wgb[(int64_t)m * W3 + w] : -1;
    lcnt[w] = 0;
  }
  __syncthreads();
  const SubcellCfg c = A.c;
  for (int j = beg + threadIdx.x; j < end; j += blockDim.x) {
    const float4 p = A.spos[j];
    const float q[3] = {p.x, p.y, p.z};
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (A.list_vel || A.list_vq16) v = A.svel[j];
    const uint32_t clr = A.list_clr ? load_u32(A.P.clr, __float_as_uint(p.w), A.P.clr_off, A.P.clr_stride) : 0u;
    int lo[3], hi[3];
    subcell_range(q, c.r, c.lgsc, lo, hi);
    if (kCorners == 2) {
      int slot[8];
      bool act[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int x = lo[0] + (e & 1) * c.sc, y = lo[1] + ((e >> 1) & 1) * c.sc, z = lo[2] + (e >> 2) * c.sc;
        const int w = (x <= hi[0] && y <= hi[1] && z <= hi[2]) ? B.window_of(x, y, z, c.lgsc) : -1;
        const int base = w >= 0 ? lbase[w] : -1;
        act[e] = base >= 0;
        slot[e] = act[e] ? base + atomicAdd(&lcnt[w], 1) : 0;
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        if (!act[e]) continue;
        put_entry(A, slot[e], q, v.x, v.y, v.z, clr);
      }
      continue;
    }
    if (kCorners == 3 && hi[0] - lo[0] <= 2 * c.sc && hi[1] - lo[1] <= 2 * c.sc && hi[2] - lo[2] <= 2 * c.sc) {
      int slot[27];
      unsigned act = 0u, direct = 0u;  // bit e: candidate e takes a window slot / a one-by-one slot
#pragma unroll
      for (int e = 0; e < 27; ++e) {
        const int x = lo[0] + (e % 3) * c.sc, y = lo[1] + ((e / 3) % 3) * c.sc, z = lo[2] + (e / 9) * c.sc;
        slot[e] = 0;
        if (x > hi[0] || y > hi[1] || z > hi[2]) continue;
        const int w = B.window_of(x, y, z, c.lgsc);
        if (w < 0) {
          direct |= 1u << e;
          continue;
        }
        const int base = lbase[w];
        if (base < 0) continue;
        act |= 1u << e;
        slot[e] = base + atomicAdd(&lcnt[w], 1);
      }
#pragma unroll
      for (int e = 0; e < 27; ++e)
        if ((act >> e) & 1u) put_entry(A, slot[e], q, v.x, v.y, v.z, clr);
      for (unsigned d = direct; d; d &= d - 1u) {
        const int e = __ffs(d) - 1;
        const int x = lo[0] + (e % 3) * c.sc, y = lo[1] + ((e / 3) % 3) * c.sc, z = lo[2] + (e / 9) * c.sc;
        const int64_t g = direct_subcell(A, x, y, z, B.ox, B.oy, B.oz);
        if (g < 0) continue;
        put_entry(A, (int64_t)A.off[g] + A.cnt_in[g] + atomicAdd(&A.ocur[g], 1), q, v.x, v.y, v.z, clr);
      }
      continue;
    }
    for (int x = lo[0]; x <= hi[0]; x += c.sc)
      for (int y = lo[1]; y <= hi[1]; y += c.sc)
        for (int z = lo[2]; z <= hi[2]; z += c.sc) {
          const int w = B.window_of(x, y, z, c.lgsc);
          int64_t k;
          if (w >= 0) {
            const int base = lbase[w];
            if (base < 0) continue;
            k = (int64_t)base + atomicAdd(&lcnt[w], 1);
          } else {
            const int64_t g = direct_subcell(A, x, y, z, B.ox, B.oy, B.oz);
            if (g < 0) continue;
            k = (int64_t)A.off[g] + A.cnt_in[g] + atomicAdd(&A.ocur[g], 1);
          }
          put_entry(A, k, q, v.x, v.y, v.z, clr);
        }
  }
}

template <int kD, int kCorners, bool kPersist>
__global__ void __launch_bounds__(256) k_bin_insert(BinArgs A, int nmap) {
  extern __shared__ unsigned long long smem64[];
  __shared__ int nb[27];
  for_each_brick<kPersist>((int64_t)nmap + kOrphanItems, A.L.queue, [&](int m) {
    if (m < nmap) bin_insert_brick<kD, kCorners>(A, m, smem64, nb);
    else bin_points_body<true>(A, A.start + nmap, m - nmap, kOrphanItems);
  });
}

int bin_uses_class_bins(int D) { return D <= 2 ? 1 : 0; }

// D = 0: no class bins
size_t bin_count_smem(int W, int D, bool vel) {
  const int W3 = W * W * W;
  const int NB = D > 0 && bin_uses_class_bins(D) ? D * D * D * 64 : 0;
  return (vel ? (size_t)(W3 + NB) * 24 : 0) + (size_t)(W3 + NB) * 4 + 8;
}

// Threads per brick workgroup.  The per-brick kernels stride every loop by blockDim, so any
// multiple of 64 is valid: as many waves as the average brick has items for (one wave for
// sparse bricks, where the time goes to each brick's chain of dependent loads and one-wave
// workgroups keep four times as many bricks in flight; up to four).  unit scales the items to
// lanes (a point of the count pass, an entry of the gather).  A launch shape that names the
// block size overrides this.
static inline int brick_threads(const LaunchCfg& L, int64_t items, int64_t bricks, int64_t unit) {
  if (L.block) return L.block;
  const int64_t per = items / (bricks > 0 ? bricks : 1) * 64 / unit;
  return per <= 64 ? 64 : (per <= 128 ? 128 : 256);
}

// Grid form: one workgroup per brick.  Persistent form: as many workgroups as fit on the device
// at once (at most one per kBrickChunk bricks), each taking bricks from the queue, which is
// cleared first by a fill kernel (so the pair also records into a graph).
template <typename... KArgs, typename... Args>
static void launch_bricks(void (*grid_k)(KArgs...), void (*pers_k)(KArgs...), const LaunchCfg& L, int bs, size_t sm,
                          int64_t items, hipStream_t s, Args... args) {
  if (L.persist && L.queue) {
    fill_u32((uint32_t*)L.queue, 1, 0u, s);
    int per = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, pers_k, bs, sm) != hipSuccess || per < 1) per = 1;
    int64_t g = (int64_t)per * (L.cus > 0 ? L.cus : 1);
    const int64_t need = (items + kBrickChunk - 1) / kBrickChunk;
    if (g > need) g = need;
    if (g < 1) g = 1;
    hipLaunchKernelGGL(pers_k, dim3((unsigned)g), dim3(bs), sm, s, args...);
  } else {  // room for for_each_brick's XCD-aware order (surplus items skip)
    hipLaunchKernelGGL(grid_k, item_grid(xcd_items(items)), dim3(bs), sm, s, args...);
  }
}

void bin_count(const BinArgs& A, int nmap, hipStream_t s) {
  if (nmap <= 0) return;
  const int bs = brick_threads(A.L, A.P.n, nmap, 64);
  // class bins pay off when many points share a brick; sparse bricks count pairs directly
  const int bins = bs == 256 ? bin_uses_class_bins(A.c.D) : 0;
  const size_t sm = bin_count_smem(A.W, bins ? A.c.D : 0, A.vsum != nullptr && A.P.vel != nullptr);
  if (A.c.D == 2)
    launch_bricks(k_bin_count<2, false>, k_bin_count<2, true>, A.L, bs, sm, (int64_t)nmap + kOrphanItems, s, A, nmap,
                  bins);  // gathers too
  else
    launch_bricks(k_bin_count<0, false>, k_bin_count<0, true>, A.L, bs, sm, (int64_t)nmap + kOrphanItems, s, A, nmap,
                  bins);
}

void bin_insert(const BinArgs& A, int nmap, hipStream_t s) {
  if (nmap <= 0) return;
  const size_t sm = (size_t)2 * A.W * A.W * A.W * 4 + 8;
  const int bs = brick_threads(A.L, A.P.n, nmap, 64);
  const bool two = 2.0 * (double)A.c.r + 2.0 <= (double)A.c.sc;
  const bool three = (double)A.c.r + 1.0 <= (double)A.c.sc;
#define GVDBX_INSERT(D, K) \
  launch_bricks(k_bin_insert<D, K, false>, k_bin_insert<D, K, true>, A.L, bs, sm, (int64_t)nmap + kOrphanItems, s, A, \
                nmap)
  if (A.c.D == 2) {
    if (two) GVDBX_INSERT(2, 2);
    else if (three) GVDBX_INSERT(2, 3);
    else GVDBX_INSERT(2, 0);
  } else {
    if (two) GVDBX_INSERT(0, 2);
    else if (three) GVDBX_INSERT(0, 3);
    else GVDBX_INSERT(0, 0);
  }
#undef GVDBX_INSERT
}

__global__ void k_sum_counts(const int* __restrict__ a, const int* __restrict__ b, int* __restrict__ out, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}

void sum_counts(const int* a, const int* b, int* out, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sum_counts, dim3(nblocks(n, 256)), dim3(256), 0, s, a, b, out, n);
}

// Sorted copies of the points (coalesced reads for the per-brick kernels).  Neighbouring
// bricks' points share cache lines of the caller's arrays, so consecutive blocks of sorted
// points go to one XCD (xcd_item), whose L2 then fetches each line once.
__global__ void k_gather_sorted(PointsView P, const uint32_t* __restrict__ svals, int64_t n, float4* __restrict__ spos,
                                float4* __restrict__ svel, int64_t blocks) {
  const int64_t blk = xcd_item(blocks);
  if (blk >= blocks) return;
  const int64_t j = blk * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = svals[j];
  const float3 p = load_f3(P.pos, i, P.pos_off, P.pos_stride);
  spos[j] = make_float4(p.x + P.tx, p.y + P.ty, p.z + P.tz, __uint_as_float(i));
  if (svel) {
    const float3 v = load_f3(P.vel, i, P.vel_off, P.vel_stride);
    svel[j] = make_float4(v.x, v.y, v.z, 0.f);
  }
}

void gather_sorted(const PointsView& P, const uint32_t* svals, int64_t n, float4* spos, float4* svel, hipStream_t s) {
  if (n <= 0) return;
  const int64_t blocks = nblocks(n, 256);
  hipLaunchKernelGGL(k_gather_sorted, dim3((unsigned)xcd_items(blocks)), dim3(256), 0, s, P, svals, n, spos, svel,
                     blocks);
}

// Walks the bricks whose window reaches subcell g = n*D^3 + l: per axis the brick itself and,
// for l on the first/last layer, the neighbour below/above (window cell l - delta*D).
// The subcell offsets (exclusive scan of cnt) come from the two passes themselves: the totals
// pass writes each 256-subcell block's sum to tile[block], one workgroup scans those, and the
// bases pass adds the block's prefix to a scan of its own counts (cnt = the counts, out = off).
template <bool kBases, int kD>
__global__ void __launch_bounds__(256) k_cell_walk(BinArgs A, int64_t nsc, int* __restrict__ cnt,
                                                   int* __restrict__ tile, int* __restrict__ off_out, int64_t blocks) {
  // consecutive blocks (neighbouring bricks, whose window counts each walk reads) on one XCD
  const int64_t blk = xcd_item(blocks);
  if (blk >= blocks) return;
  const int64_t g = blk * blockDim.x + threadIdx.x;
  const bool live = g < nsc;
  const int D = kD ? kD : A.c.D, D3 = D * D * D, W = D + 2, W3 = W * W * W;
  // every lane reaches the one block scan of each pass (it holds barriers): the bases pass scans
  // first, the totals pass after its walk, both outside any lane-dependent branch
  int base = 0;
  if (kBases) {
    int tot;
    const int c0 = live ? cnt[g] : 0;
    base = tile[blk] + block_exclusive_scan_256(c0, &tot);
    if (g == nsc - 1) off_out[nsc] = base + c0;  // the list length closes the offsets
  }
  int p = 0;
  unsigned long long vs[3] = {0ull, 0ull, 0ull};
  const bool vel = !kBases && A.vsum != nullptr && A.P.vel != nullptr;
  if (live) {
    if (kBases) {
      off_out[g] = base;
      A.ocur[g] = 0;  // the insert's one-by-one cursor starts at zero
    }
    const int64_t n = g / D3;
    const int l = (int)(g - n * D3);
    const int lx = l % D, ly = (l / D) % D, lz = l / (D * D);
    for (int dz = (lz == 0 ? -1 : 0); dz <= (lz == D - 1 ? 1 : 0); ++dz)
      for (int dy = (ly == 0 ? -1 : 0); dy <= (ly == D - 1 ? 1 : 0); ++dy)
        for (int dx = (lx == 0 ? -1 : 0); dx <= (lx == D - 1 ? 1 : 0); ++dx) {
          const int b = A.nbm[n * 27 + (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
          if (b < 0) continue;
          const int w = ((lz - dz * D + 1) * W + (ly - dy * D + 1)) * W + (lx - dx * D + 1);
          int* cell = A.wgb + (int64_t)b * W3 + w;
          const int c = *cell;
          // the insert reads a cell's base only where its brick places entries (c > 0)
          if (kBases && c > 0) *cell = base + p;
          p += c;
          if (vel && c > 0) {
            const unsigned long long* v = A.wvs + ((int64_t)b * W3 + w) * 3;
            vs[0] += v[0];
            vs[1] += v[1];
            vs[2] += v[2];
          }
        }
  }
  if (!kBases) {
    int total = 0;
    if (live) {
      A.cnt_in[g] = p;
      total = p + A.ocnt[g];
      cnt[g] = total;
      if (vel && p > 0) {  // the pairs placed one by one added theirs with atomics before
        A.vsum[3 * g + 0] += vs[0];
        A.vsum[3 * g + 1] += vs[1];
        A.vsum[3 * g + 2] += vs[2];
      }
    }
    int tot;
    (void)block_exclusive_scan_256(total, &tot);
    if (threadIdx.x == 0) tile[blk] = tot;
  }
}

void cell_totals(const BinArgs& A, int64_t nsc, int* cnt, int* tile, hipStream_t s) {
  if (nsc <= 0) return;
  const int64_t blocks = nblocks(nsc, 256);
  const dim3 grid((unsigned)xcd_items(blocks));
  if (A.c.D == 2)
    hipLaunchKernelGGL((k_cell_walk<false, 2>), grid, dim3(256), 0, s, A, nsc, cnt, tile, (int*)nullptr, blocks);
  else
    hipLaunchKernelGGL((k_cell_walk<false, 0>), grid, dim3(256), 0, s, A, nsc, cnt, tile, (int*)nullptr, blocks);
}

void cell_bases(const BinArgs& A, int64_t nsc, int* cnt, int* tile, int* off, hipStream_t s) {
  if (nsc <= 0) return;
  const int64_t blocks = nblocks(nsc, 256);
  const dim3 grid((unsigned)xcd_items(blocks));
  if (A.c.D == 2)
    hipLaunchKernelGGL((k_cell_walk<true, 2>), grid, dim3(256), 0, s, A, nsc, cnt, tile, off, blocks);
  else
    hipLaunchKernelGGL((k_cell_walk<true, 0>), grid, dim3(256), 0, s, A, nsc, cnt, tile, off, blocks);
}

// ------------------------------------------------------------------ per-point path
// Sorted positions [begin[0], begin[1]) (the points whose own leaf is missing) are binned one
// pair at a time with global atomics.
// ------------------------------------------------------------------ gathers

__device__ __forceinline__ int64_t atlas_index(int slot, int br, int lx, int ly, int lz) {
  return (((int64_t)slot * br + lz) * br + ly) * br + lx;
}

__device__ __forceinline__ float mean_of(const int* cnt, const unsigned long long* vsum, int64_t k, int a, int acc,
                                         float old) {
  const double den = (double)((long long)cnt[k] + acc);
  double v = (double)(long long)vsum[3 * k + a] * (1.0 / kVelScale);
  if (acc) v += (double)old;
  return den > 0 ? (float)(v / den) : 0.0f;
}

// GatherDensity without accumulation (P.cuh:696-733): val = max(0, max_p (r - |p - v|)) over
// the points listed for the voxel's subcell.  Only points within r of a voxel can raise it
// above 0, so each list entry visits the voxels of its own subcell inside its r-ball (a
// 0.01-voxel margin keeps rounding at the ball edge inclusive) and min-reduces |p-v|^2 in
// LDS.  One workgroup per brick, one wave per subcell.
//
// The apron is produced here too: every point within r of an apron voxel g lies in the list
// of the brick subcell next to g (the list box [q-r, q+1+r] reaches it), so extending the
// ball push of boundary subcells by the apron width yields exactly the value the neighbour
// brick computes for g.  Apron voxels whose neighbour brick exists are written with it; the
// others keep waiting for UpdateApron's boundary value.
//
// kVel: the velocity means of GatherVelocity (non-accumulating, k_velocity_brick's rule) are
// written in the same pass, as the reference computes its colour channel inside the density
// gather (P.cuh:717-731): window-cell means in LDS, three more stores per voxel.
//
// kCorner (fixed geometry with the apron): each point corner takes max(0, r - sqrt(d2)) itself
// and max-reduces its bits in LDS (non-negative floats order as their bits), so the output loop
// only copies.  f(d) = max(0, r - sqrtf(d)) is monotone non-increasing with correctly rounded
// sqrtf and subtraction, so max_p f(d_p) = f(min_p d_p): the same bits as the per-voxel form,
// with one square root per point corner instead of one per voxel (sparse bricks have far fewer
// corners than voxels; dense bricks the other way round).

// The fixed geometry's padded voxels (8^3 brick, apron 1, 2^3 subcells): entry t of the 10^3
// padded brick holds in bits 0-4 the neighbour index (oz+1)*9 + (oy+1)*3 + (ox+1) of the brick
// the voxel lies in (13 = this one), in bits 8-13 its window cell ((wz*4 + wy)*4 + wx, each
// w = floor(coord / 4) + 1), so the output loop decodes no coordinates.
template <int kRes, int kApron, int kD>
struct PadTable {
  static constexpr int kBw = kRes + 2 * kApron, kVox = kBw * kBw * kBw, kW = kD + 2, kSc = kRes / kD;
  unsigned short v[kVox];
  static constexpr int side(int x) { return x < 0 ? 0 : (x >= kRes ? 2 : 1); }
  static constexpr int cell(int x) { return (x < 0 ? -1 : x / kSc) + 1; }  // apron <= subcell size
  constexpr PadTable() : v() {
    for (int t = 0; t < kVox; ++t) {
      const int x = t % kBw - kApron, y = (t / kBw) % kBw - kApron, z = t / (kBw * kBw) - kApron;
      const int nbi = side(z) * 9 + side(y) * 3 + side(x);
      const int w = (cell(z) * kW + cell(y)) * kW + cell(x);
      v[t] = (unsigned short)(nbi | (w << 8));
    }
  }
};
__constant__ __align__(16) PadTable<8, 1, 2> kPad812 = PadTable<8, 1, 2>();

template <int kRes, int kApron, int kD, bool kWithApron, bool kVel, bool kCorner>
__device__ __forceinline__ void density_brick_body(const SubcellCfg& c, int64_t m, int64_t nsc, int64_t len,
                                                   const int* __restrict__ map_leaf, const int* __restrict__ off,
                                                   const float* __restrict__ list_pos,
                                                   const int* __restrict__ leaf_slot,
                                                   const gvdbx_node* __restrict__ nodes0,
                                                   const int4* __restrict__ slot_pos, int apron_rt,
                                                   const int* __restrict__ nbr, float* __restrict__ chan,
                                                   const VelOut& vo) {
  // LDS sized for the compile-time geometry when there is one (more bricks fit on a CU)
  constexpr int kVox = kRes ? (kRes + 2 * kApron) * (kRes + 2 * kApron) * (kRes + 2 * kApron) : kMaxPaddedVox;
  constexpr int kOff = kRes ? kD * kD * kD + 1 : 513;
  __shared__ __align__(16) unsigned int d2[kVox];
  __shared__ int soff[kOff];
  __shared__ int nb[27];
  constexpr int kMeanCells = kVel ? (kRes ? (kD + 2) * (kD + 2) * (kD + 2) : kMaxWindowCells) : 1;
  __shared__ float vmean[kMeanCells][3];
  // geometry: compile-time for the default 8^3 bricks / apron 1 / 2^3 subcells, else runtime
  const int res0 = kRes ? kRes : c.res0;
  const int apron = kRes ? kApron : apron_rt;
  const int D = kRes ? kD : c.D;
  const int sc = kRes ? kRes / kD : c.sc;
  const int a = kWithApron ? apron : 0;   // apron width computed here
  const int br = res0 + 2 * apron;        // channel brick layout
  const int bw = res0 + 2 * a;            // computed region
  const int D3 = D * D * D;
  // brick corner and slot: with the slot map (slot_pos given) brick m is slot m, one load
  int slot, gx0, gy0, gz0;
  if (slot_pos) {
    const int4 sp = slot_pos[m];
    slot = (int)m;
    gx0 = sp.x;
    gy0 = sp.y;
    gz0 = sp.z;
  } else {
    const int leaf = map_leaf[m];
    if (leaf < 0) return;
    slot = leaf_slot[leaf];
    gx0 = nodes0[leaf].pos[0];
    gy0 = nodes0[leaf].pos[1];
    gz0 = nodes0[leaf].pos[2];
  }
  const int nv = bw * bw * bw;
  const int64_t k0 = (int64_t)m * D3;
  static_assert(!kCorner || (kRes == 8 && kApron == 1 && kD == 2 && kWithApron), "kCorner: fixed geometry only");
  constexpr unsigned int kInit = kCorner ? 0u : 0x7F800000u;  // value 0, or squared distance +inf
  if constexpr (kRes && kWithApron && (kVox & 3) == 0) {  // nv == kVox: 16-byte stores
    for (int t4 = threadIdx.x; t4 < kVox / 4; t4 += blockDim.x)
      reinterpret_cast<uint4*>(d2)[t4] = make_uint4(kInit, kInit, kInit, kInit);
  } else {
    for (int t = threadIdx.x; t < nv; t += blockDim.x) d2[t] = kInit;
  }
  for (int s = threadIdx.x; s <= D3; s += blockDim.x) soff[s] = off[k0 + s];  // off[nsc] = the list length
  if (threadIdx.x < 27) nb[threadIdx.x] = (kWithApron && slot >= 0) ? nbr[(int64_t)slot * 27 + threadIdx.x] : -1;
  __syncthreads();
  const int bx = gx0 & c.cmask, by = gy0 & c.cmask, bz = gz0 & c.cmask;
  if (kVel) {  // means of the window subcells (own D^3 and the one-subcell shell)
    const int W = D + 2, W3 = W * W * W;
    for (int w = threadIdx.x; w < W3; w += blockDim.x) {
      const int sx = w % W - 1, sy = (w / W) % W - 1, sz = w / (W * W) - 1;
      const int ox = sx < 0 ? -1 : (sx >= D ? 1 : 0), oy = sy < 0 ? -1 : (sy >= D ? 1 : 0),
                oz = sz < 0 ? -1 : (sz >= D ? 1 : 0);
      const int b = (ox | oy | oz) == 0 ? (int)m : nb[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)];
      if (b < 0) {  // apron voxels without a neighbour brick: boundary value 0
#pragma unroll
        for (int q = 0; q < 3; ++q) vmean[w][q] = 0.0f;
        continue;
      }
      // slot order equals subcell-map order
      const int64_t k = (int64_t)b * D3 + ((sz - oz * D) * D + (sy - oy * D)) * D + (sx - ox * D);
#pragma unroll
      for (int q = 0; q < 3; ++q) vmean[w][q] = mean_of(vo.cnt, vo.vsum, k, q, 0, 0.0f);
    }
  }
  // Voxels v with |p - v| < r lie in the integer cube [floor(p) - R + 1, floor(p) + R], R = ceil(r);
  // voxels at distance >= r cannot raise max(0, r - |p - v|), so the cube is exact.
  const int R = (int)ceilf(c.r);
  // The brick's subcell lists are contiguous, [soff[0], soff[D3]): the lanes take its entries in
  // order and each finds its subcell by binary search over the offsets (sparse bricks fill the
  // lanes across subcells instead of running one short list per wave).
  // fixed geometry: the D3 + 1 offsets in scalar registers, the subcell by counting starts <= j
  constexpr int kFixedD3 = kRes ? kD * kD * kD : 1;
  int so[kFixedD3 + 1];
  if (kRes) {
#pragma unroll
    for (int k = 0; k <= kFixedD3; ++k) so[k] = __builtin_amdgcn_readfirstlane(soff[k]);
  }
  // each entry's position is loaded one iteration ahead, so its latency overlaps the
  // previous entry's corner updates
  const int jend = soff[D3];
  float3 gnext = make_float3(0.f, 0.f, 0.f);
  if (soff[0] + (int)threadIdx.x < jend) {
    const int64_t j0 = soff[0] + (int)threadIdx.x;
    gnext = make_float3(list_pos[3 * j0], list_pos[3 * j0 + 1], list_pos[3 * j0 + 2]);
  }
  for (int j = soff[0] + (int)threadIdx.x; j < jend; j += blockDim.x) {
    const float3 gcur = gnext;
    const int jn = j + (int)blockDim.x;
    if (jn < jend) gnext = make_float3(list_pos[3 * (int64_t)jn], list_pos[3 * (int64_t)jn + 1],
                                       list_pos[3 * (int64_t)jn + 2]);
    unsigned s = 0;
    if (kRes) {
#pragma unroll
      for (int k = 1; k < kFixedD3; ++k) s += j >= so[k] ? 1u : 0u;
    } else {
      unsigned hi = D3;  // soff[s] <= j < soff[hi]
      while (hi - s > 1) {
        const unsigned mid = (s + hi) >> 1;
        if (soff[mid] <= j) s = mid; else hi = mid;
      }
    }
    const int ix = (int)(s % (unsigned)D), iy = (int)((s / (unsigned)D) % (unsigned)D),
              iz = (int)(s / (unsigned)(D * D));
    // voxel range of the subcell, extended into the apron on the brick's boundary
    const int lx0 = ix * sc - (ix == 0 ? a : 0), lx1 = ix * sc + sc - 1 + (ix == D - 1 ? a : 0);
    const int ly0 = iy * sc - (iy == 0 ? a : 0), ly1 = iy * sc + sc - 1 + (iy == D - 1 ? a : 0);
    const int lz0 = iz * sc - (iz == 0 ? a : 0), lz1 = iz * sc + sc - 1 + (iz == D - 1 ? a : 0);
    {
      const float gx = gcur.x, gy = gcur.y, gz = gcur.z;
      const int fx = (int)gx - bx, fy = (int)gy - by, fz = (int)gz - bz;
      if (R == 1) {
        // corners fx + {0, 1} per axis: offsets and x squares once per entry, then each corner's
        // |v - p|^2 as dist2 forms it (fmaf(z, z, fmaf(y, y, x * x))), bit for bit
        const float ex[2] = {gx - (float)(bx + fx), gx - (float)(bx + fx + 1)};
        const float ey[2] = {gy - (float)(by + fy), gy - (float)(by + fy + 1)};
        const float ez[2] = {gz - (float)(bz + fz), gz - (float)(bz + fz + 1)};
        const float px[2] = {ex[0] * ex[0], ex[1] * ex[1]};
        const bool vx[2] = {fx >= lx0 && fx <= lx1, fx + 1 >= lx0 && fx + 1 <= lx1};
        const bool vy[2] = {fy >= ly0 && fy <= ly1, fy + 1 >= ly0 && fy + 1 <= ly1};
        const bool vz[2] = {fz >= lz0 && fz <= lz1, fz + 1 >= lz0 && fz + 1 <= lz1};
        const int i0 = ((fz + a) * bw + (fy + a)) * bw + (fx + a);
#pragma unroll
        for (int dz = 0; dz < 2; ++dz)
#pragma unroll
          for (int dy = 0; dy < 2; ++dy)
#pragma unroll
            for (int dx = 0; dx < 2; ++dx) {
              if (!(vx[dx] && vy[dy] && vz[dz])) continue;
              const float d = fmaf(ez[dz], ez[dz], fmaf(ey[dy], ey[dy], px[dx]));
              if (kCorner) {
                const float val = fmaxf(0.0f, c.r - sqrtf(d));
                if (val > 0.0f) atomicMax(&d2[i0 + (dz * bw + dy) * bw + dx], __float_as_uint(val));
              } else {
                atomicMin(&d2[i0 + (dz * bw + dy) * bw + dx], __float_as_uint(d));
              }
            }
      } else {
        const int x0 = max(lx0, fx - R + 1), x1 = min(lx1, fx + R);
        const int y0 = max(ly0, fy - R + 1), y1 = min(ly1, fy + R);
        const int z0 = max(lz0, fz - R + 1), z1 = min(lz1, fz + R);
        for (int z = z0; z <= z1; ++z)
          for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) {
              const float d = dist2(gx - (float)(bx + x), gy - (float)(by + y), gz - (float)(bz + z));
              unsigned int* cell = &d2[((z + a) * bw + (y + a)) * bw + (x + a)];
              if (kCorner) {
                const float val = fmaxf(0.0f, c.r - sqrtf(d));
                if (val > 0.0f) atomicMax(cell, __float_as_uint(val));
              } else {
                atomicMin(cell, __float_as_uint(d));
              }
            }
      }
    }
  }
  __syncthreads();
  if (slot < 0) return;
  float* bptr = chan + (int64_t)slot * br * br * br;
  if constexpr (kRes == 8 && kApron == 1 && kD == 2 && kWithApron) {
    // fixed geometry with the apron: the computed region is the whole padded brick, so voxels
    // t..t+3 are consecutive in LDS and in the channel (16-byte loads and stores); each voxel's
    // neighbour index and window cell come from the padded-voxel table
    // the missing neighbours as a bit mask (wave-uniform): their apron voxels get 0
    const int ln = (int)(threadIdx.x & 63);
    const unsigned int miss = (unsigned int)__ballot(ln < 27 && nb[ln < 27 ? ln : 0] < 0);
    const uint2* tab = reinterpret_cast<const uint2*>(kPad812.v);
    for (int t4 = threadIdx.x; t4 < kVox / 4; t4 += blockDim.x) {
      const uint4 u4 = reinterpret_cast<const uint4*>(d2)[t4];
      const uint2 tb = tab[t4];
      const unsigned int u[4] = {u4.x, u4.y, u4.z, u4.w};
      const unsigned int ent[4] = {tb.x & 0xFFFFu, tb.x >> 16, tb.y & 0xFFFFu, tb.y >> 16};
      float val[4], vm[3][4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const bool missing = (miss >> (ent[e] & 31u)) & 1u;
        const float f = __uint_as_float(u[e]);
        // per voxel: f = min d2 (+inf when no point reached it, giving max(0, -inf) = 0)
        val[e] = missing ? 0.0f : (kCorner ? f : fmaxf(0.0f, c.r - sqrtf(f)));
        if (kVel) {
          const unsigned int w = ent[e] >> 8;
#pragma unroll
          for (int q = 0; q < 3; ++q) vm[q][e] = vmean[w][q];  // 0 for missing neighbours
        }
      }
      // streaming stores: nothing in the step reads the channel bricks back
      store_stream4(bptr + 4 * t4, val[0], val[1], val[2], val[3]);
      if (kVel) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
          store_stream4(vo.ch[q] + (int64_t)slot * kVox + 4 * t4, vm[q][0], vm[q][1], vm[q][2], vm[q][3]);
      }
    }
    return;
  }
  for (int t = threadIdx.x; t < nv; t += blockDim.x) {
    const int zy = t / bw;
    const int x = t - zy * bw - a, z = zy / bw - a, y = zy - (zy / bw) * bw - a;
    bool missing = false;  // apron voxel without a neighbour brick: boundary value 0
    if (kWithApron) {
      const int ox = x < 0 ? -1 : (x >= res0 ? 1 : 0), oy = y < 0 ? -1 : (y >= res0 ? 1 : 0),
                oz = z < 0 ? -1 : (z >= res0 ? 1 : 0);
      missing = (ox | oy | oz) != 0 && nb[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)] < 0;
    }
    const unsigned int u = d2[t];
    const float val = (missing || u == 0x7F800000u) ? 0.0f : fmaxf(0.0f, c.r - sqrtf(__uint_as_float(u)));
    // the computed region is the whole padded brick when the apron is computed here
    const int64_t vi = a == apron ? t : ((z + apron) * br + (y + apron)) * br + (x + apron);
    bptr[vi] = val;
    if (kVel) {
      const int lgsc = kRes ? 31 - __builtin_clz(kRes / kD) : c.lgsc;
      const int W = D + 2;
      const int w = (((z >> lgsc) + 1) * W + ((y >> lgsc) + 1)) * W + ((x >> lgsc) + 1);
      const int64_t gi = (int64_t)slot * br * br * br + vi;
#pragma unroll
      for (int q = 0; q < 3; ++q) vo.ch[q][gi] = missing ? 0.0f : vmean[w][q];
    }
  }
}

template <int kRes, int kApron, int kD, bool kWithApron, bool kVel, bool kCorner, bool kPersist>
__global__ void __launch_bounds__(256) k_density_brick(SubcellCfg c, int nmap, int64_t nsc, int64_t len,
                                                      const int* __restrict__ map_leaf, const int* __restrict__ off,
                                                      const float* __restrict__ list_pos,
                                                      const int* __restrict__ leaf_slot,
                                                      const gvdbx_node* __restrict__ nodes0,
                                                      const int4* __restrict__ slot_pos, int apron_rt,
                                                      const int* __restrict__ nbr, float* __restrict__ chan,
                                                      VelOut vo, int* queue) {
  for_each_brick<kPersist>(nmap, queue, [&](int m) {
    density_brick_body<kRes, kApron, kD, kWithApron, kVel, kCorner>(c, m, nsc, len, map_leaf, off, list_pos, leaf_slot,
                                                           nodes0, slot_pos, apron_rt, nbr, chan, vo);
  });
}

void density_brick(SubcellCfg c, int nmap, int64_t nsc, int64_t len, const int* map_leaf, const int* off,
                   const float* list_pos, const int* leaf_slot, const gvdbx_node* nodes0, const int4* slot_pos,
                   int apron, const int* nbr, float* chan, const VelOut* vel, const LaunchCfg& L, hipStream_t s) {
  if (nmap <= 0) return;
  // a wave per 512 entries of the average brick: with fewer lanes per brick a CU holds more
  // bricks, whose set-up loads, entry loop and channel writes then overlap (4M SPH points, 900
  // entries per brick: 128 threads 126 us, 256 threads 137 us)
  const int bs = brick_threads(L, len, nmap, 512);
  const bool fixed = c.res0 == 8 && c.D == 2 && apron == 1;
  const VelOut vo = vel ? *vel : VelOut{nullptr, nullptr, {nullptr, nullptr, nullptr}};
  // square roots per point corner when the bricks hold few list entries (an entry visits about
  // 2 corners at r = 1 and subcells of 4, a brick has 1000 padded voxels): the same bits either way
  const bool corner = len < (int64_t)256 * nmap;
#define GVDBX_DENSITY(R, A, D, AP, V, C)                                                                           \
  launch_bricks(k_density_brick<R, A, D, AP, V, C, false>, k_density_brick<R, A, D, AP, V, C, true>, L, bs, 0, nmap, \
                s, c, nmap, nsc, len, map_leaf, off, list_pos, leaf_slot, nodes0, slot_pos, apron, nbr, chan, vo,    \
                L.queue)
  if (vel) {  // velocity means need the apron pass (a neighbour table)
    if (fixed && corner) GVDBX_DENSITY(8, 1, 2, true, true, true);
    else if (fixed) GVDBX_DENSITY(8, 1, 2, true, true, false);
    else GVDBX_DENSITY(0, 0, 0, true, true, false);
  } else if (fixed && nbr) {
    if (corner) GVDBX_DENSITY(8, 1, 2, true, false, true);
    else GVDBX_DENSITY(8, 1, 2, true, false, false);
  } else if (fixed) {
    GVDBX_DENSITY(8, 1, 2, false, false, false);
  } else if (nbr) {
    GVDBX_DENSITY(0, 0, 0, true, false, false);
  } else {
    GVDBX_DENSITY(0, 0, 0, false, false, false);
  }
#undef GVDBX_DENSITY
}

// Velocity means from the fixed-point sums: (accum ? old : 0) + sum, over (cnt + accum); the
// colour rule of P.cuh:717-731 for float3; an empty subcell without accumulation gives 0.
// One workgroup per brick.  Without accumulation the apron voxels whose neighbour exists get
// the mean of the neighbour subcell holding them (slot order equals subcell-map order).

template <int kRes, int kApron, int kD>
__global__ void __launch_bounds__(256) k_velocity_brick(SubcellCfg c, int nmap, const int* __restrict__ map_leaf,
                                                       const int* __restrict__ cnt,
                                                       const unsigned long long* __restrict__ vsum,
                                                       const int* __restrict__ leaf_slot, int apron_rt,
                                                       const int* __restrict__ nbr, float* __restrict__ cx,
                                                       float* __restrict__ cy, float* __restrict__ cz, int accum) {
  __shared__ int nb[27];
  __shared__ float mean[kMaxWindowCells][3];
  const int res0 = kRes ? kRes : c.res0;
  const int apron = kRes ? kApron : apron_rt;
  const int D = kRes ? kD : c.D;
  const int lgsc = kRes ? (kD == 2 ? 2 : c.lgsc) : c.lgsc;
  const int64_t m = item_index();
  if (m >= nmap) return;
  const int leaf = map_leaf[m];
  if (leaf < 0) return;
  const int slot = leaf_slot[leaf];
  if (slot < 0) return;
  const int D3 = D * D * D;
  const int a = (nbr && !accum) ? apron : 0;
  if (threadIdx.x < 27) nb[threadIdx.x] = a ? nbr[(int64_t)slot * 27 + threadIdx.x] : -1;
  __syncthreads();
  // means of the window subcells (own D^3 plus the one-subcell shell when the apron is written)
  const int W = D + 2, W3 = W * W * W;
  if (!accum) {
    for (int w = threadIdx.x; w < W3; w += blockDim.x) {
      const int sx = w % W - 1, sy = (w / W) % W - 1, sz = w / (W * W) - 1;
      const int ox = sx < 0 ? -1 : (sx >= D ? 1 : 0), oy = sy < 0 ? -1 : (sy >= D ? 1 : 0),
                oz = sz < 0 ? -1 : (sz >= D ? 1 : 0);
      const int b = (ox | oy | oz) == 0 ? m : nb[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)];
      if (b < 0) continue;
      const int64_t k = (int64_t)b * D3 + ((sz - oz * D) * D + (sy - oy * D)) * D + (sx - ox * D);
      for (int q = 0; q < 3; ++q) mean[w][q] = mean_of(cnt, vsum, k, q, 0, 0.0f);
    }
    __syncthreads();
  }
  const int br = res0 + 2 * apron;
  const int bw = res0 + 2 * a;  // the region written
  const int nv = bw * bw * bw;
  const int64_t base = (int64_t)slot * br * br * br;
  float* ch[3] = {cx, cy, cz};
  for (int t = threadIdx.x; t < nv; t += blockDim.x) {
    const int zy = t / bw;
    const int x = t - zy * bw - a, z = zy / bw - a, y = zy - (zy / bw) * bw - a;
    const int ox = x < 0 ? -1 : (x >= res0 ? 1 : 0), oy = y < 0 ? -1 : (y >= res0 ? 1 : 0),
              oz = z < 0 ? -1 : (z >= res0 ? 1 : 0);
    // apron voxel without a neighbour brick: the boundary value 0 (without accumulation)
    const bool missing = (ox | oy | oz) != 0 && nb[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)] < 0;
    if (missing && accum) continue;
    const int64_t vi = base + ((z + apron) * br + (y + apron)) * br + (x + apron);
    if (!accum) {
      const int w = (((z >> lgsc) + 1) * W + ((y >> lgsc) + 1)) * W + ((x >> lgsc) + 1);
#pragma unroll
      for (int q = 0; q < 3; ++q) ch[q][vi] = missing ? 0.0f : mean[w][q];
    } else {
      const int64_t k = (int64_t)m * D3 + ((z >> lgsc) * D + (y >> lgsc)) * D + (x >> lgsc);
#pragma unroll
      for (int q = 0; q < 3; ++q) ch[q][vi] = mean_of(cnt, vsum, k, q, 1, ch[q][vi]);
    }
  }
}

void velocity_brick(SubcellCfg c, int nmap, const int* map_leaf, const int* cnt, const unsigned long long* vsum,
                    const int* leaf_slot, int apron, const int* nbr, float* cx, float* cy, float* cz, int accum,
                    hipStream_t s) {
  if (nmap <= 0) return;
  if (c.res0 == 8 && c.D == 2 && apron == 1)
    hipLaunchKernelGGL((k_velocity_brick<8, 1, 2>), item_grid(nmap), dim3(256), 0, s, c, nmap, map_leaf, cnt, vsum, leaf_slot,
                       apron, nbr, cx, cy, cz, accum);
  else
    hipLaunchKernelGGL((k_velocity_brick<0, 0, 0>), item_grid(nmap), dim3(256), 0, s, c, nmap, map_leaf, cnt, vsum, leaf_slot,
                       apron, nbr, cx, cy, cz, accum);
}

// ------------------------------------------------------------------ colour means
// Colour rule of GatherDensity / GatherLevelSet (P.cuh:713-731, 752-770), per component q of
// the uchar4 words: (start_q + sum of the listed colours' q) / (cnt + accum), start = the old
// voxel with accumulation, else (0, 0, 0, 255) -- the reference's (0, 0, 0, 1) in units of
// 1/255.  Sums are exact integers and the quotient is floored and saturated at 255 (the
// reference's uchar(c * 255) truncation of the same ratio); an empty subcell without
// accumulation gives 0.

__device__ __forceinline__ uint32_t colour_mean(const unsigned long long* s, unsigned long long den, uint32_t start) {
  uint32_t out = 0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const unsigned long long num = s[q] + ((start >> (8 * q)) & 255u);
    const unsigned long long v = den ? num / den : 0ull;
    out |= (uint32_t)(v < 255ull ? v : 255ull) << (8 * q);
  }
  return out;
}

// one thread per subcell: its list segment [off, off + cnt) summed component by component
__global__ void __launch_bounds__(256) k_colour_sums(int64_t nsc, const int* __restrict__ off,
                                                    const int* __restrict__ cnt,
                                                    const uint32_t* __restrict__ list_clr,
                                                    unsigned long long* __restrict__ csum,
                                                    uint32_t* __restrict__ cmean) {
  const int64_t g = item_index() * blockDim.x + threadIdx.x;
  if (g >= nsc) return;
  const int b = off[g], n = cnt[g];
  unsigned long long sum[4] = {0ull, 0ull, 0ull, 0ull};
  for (int j = b; j < b + n; ++j) {
    const uint32_t w = list_clr[j];
#pragma unroll
    for (int q = 0; q < 4; ++q) sum[q] += (w >> (8 * q)) & 255u;
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) csum[4 * g + q] = sum[q];
  cmean[g] = colour_mean(sum, (unsigned long long)n, 0xFF000000u);
}

void colour_sums(int64_t nsc, const int* off, const int* cnt, const uint32_t* list_clr, unsigned long long* csum,
                 uint32_t* cmean, hipStream_t s) {
  if (nsc <= 0) return;
  hipLaunchKernelGGL(k_colour_sums, item_grid(nblocks(nsc, 256)), dim3(256), 0, s, nsc, off, cnt, list_clr, csum,
                     cmean);
}

// One workgroup per brick (subcell-map index m; linked leaves have slot == m, so neighbour
// slots index the subcell arrays as well).
__global__ void __launch_bounds__(256) k_colour_brick(SubcellCfg c, int nmap, const int* __restrict__ map_leaf,
                                                     const int* __restrict__ cnt,
                                                     const unsigned long long* __restrict__ csum,
                                                     const uint32_t* __restrict__ cmean,
                                                     const int* __restrict__ leaf_slot, int apron,
                                                     const int* __restrict__ nbr, uint32_t* __restrict__ chan,
                                                     int accum) {
  __shared__ int nb[27];
  const int64_t m = item_index();
  if (m >= nmap) return;
  const int leaf = map_leaf[m];
  if (leaf < 0) return;
  const int slot = leaf_slot[leaf];
  if (slot < 0) return;
  const int res0 = c.res0, D = c.D, D3 = D * D * D, lgsc = c.lgsc;
  const int a = (nbr && !accum) ? apron : 0;
  if (threadIdx.x < 27) nb[threadIdx.x] = a ? nbr[(int64_t)slot * 27 + threadIdx.x] : -1;
  __syncthreads();
  const int br = res0 + 2 * apron, bw = res0 + 2 * a, nv = bw * bw * bw;
  const int64_t base = (int64_t)slot * br * br * br;
  for (int t = threadIdx.x; t < nv; t += blockDim.x) {
    const int zy = t / bw;
    const int x = t - zy * bw - a, z = zy / bw - a, y = zy - (zy / bw) * bw - a;
    const int ox = x < 0 ? -1 : (x >= res0 ? 1 : 0), oy = y < 0 ? -1 : (y >= res0 ? 1 : 0),
              oz = z < 0 ? -1 : (z >= res0 ? 1 : 0);
    const int b = (ox | oy | oz) == 0 ? (int)m : nb[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)];
    const int64_t vi = base + ((int64_t)(z + apron) * br + (y + apron)) * br + (x + apron);
    if (b < 0) {  // apron voxel without a neighbour brick: the boundary value 0
      chan[vi] = 0u;
      continue;
    }
    const int lx = x - ox * res0, ly = y - oy * res0, lz = z - oz * res0;
    const int64_t k = (int64_t)b * D3 + ((lz >> lgsc) * D + (ly >> lgsc)) * D + (lx >> lgsc);
    if (!accum) {
      chan[vi] = cmean[k];
    } else {
      chan[vi] = colour_mean(csum + 4 * k, (unsigned long long)cnt[k] + 1ull, chan[vi]);
    }
  }
}

void colour_brick(SubcellCfg c, int nmap, const int* map_leaf, const int* cnt, const unsigned long long* csum,
                  const uint32_t* cmean, const int* leaf_slot, int apron, const int* nbr, uint32_t* chan, int accum,
                  hipStream_t s) {
  if (nmap <= 0) return;
  hipLaunchKernelGGL(k_colour_brick, item_grid(nmap), dim3(256), 0, s, c, nmap, map_leaf, cnt, csum, cmean, leaf_slot,
                     apron, nbr, chan, accum);
}

// Boundary value into the apron voxels that have no neighbour brick (the part of
// UpdateApron left once the neighbour copies are current).  One wave per brick walks only the
// apron voxels, from a table of (voxel index | neighbour direction << 27) entries.
__global__ void __launch_bounds__(64) k_apron_bound(int64_t nslots, const int* __restrict__ nbr,
                                                   const uint32_t* __restrict__ tab, int ntab, int bv,
                                                   ApronChannels ch) {
  __shared__ int nb[27];
  const int64_t s = item_index();
  if (s >= nslots) return;
  const int lane = threadIdx.x;
  const int v = lane < 27 ? nbr[s * 27 + lane] : 0;
  if (lane < 27) nb[lane] = v;
  if (__ballot(lane < 27 && v < 0) == 0) return;  // every neighbour exists (wave-uniform)
  __syncthreads();
  for (int u = lane; u < ntab; u += 64) {
    const uint32_t e = tab[u];
    if (nb[e >> 27] >= 0) continue;
    const int64_t dst = s * (int64_t)bv + (int64_t)(e & 0x7FFFFFFu);
    for (int c = 0; c < ch.n; ++c) ch.ptr[c][dst] = ch.bound[c];
  }
}

void apron_table(int res0, int apron, std::vector<uint32_t>& tab) {
  const int br = res0 + 2 * apron;
  tab.clear();
  for (int z = 0; z < br; ++z)
    for (int y = 0; y < br; ++y)
      for (int x = 0; x < br; ++x) {
        const int gx = x - apron, gy = y - apron, gz = z - apron;
        const int ox = gx < 0 ? -1 : (gx >= res0 ? 1 : 0), oy = gy < 0 ? -1 : (gy >= res0 ? 1 : 0),
                  oz = gz < 0 ? -1 : (gz >= res0 ? 1 : 0);
        if ((ox | oy | oz) == 0) continue;
        const uint32_t k = (uint32_t)((oz + 1) * 9 + (oy + 1) * 3 + (ox + 1));
        tab.push_back((uint32_t)((z * br + y) * br + x) | (k << 27));
      }
}

void apron_bound(int64_t nslots, const int* nbr, const uint32_t* tab, int ntab, int res0, int apron,
                 const ApronChannels& ch, hipStream_t s) {
  if (nslots <= 0 || apron <= 0 || ch.n <= 0) return;
  const int br = res0 + 2 * apron;
  hipLaunchKernelGGL(k_apron_bound, item_grid(nslots), dim3(64), 0, s, nslots, nbr, tab, ntab, br * br * br, ch);
}

// Per-subcell form of the velocity means, for bricks with more than 512 subcells.
__global__ void k_velocity_subcell(SubcellCfg c, int64_t nsc, const int* __restrict__ sc_pos, const int* __restrict__ sc_nid,
                                   const int* __restrict__ cnt, const unsigned long long* __restrict__ vsum,
                                   const int* __restrict__ leaf_slot, int apron, float* __restrict__ cx,
                                   float* __restrict__ cy, float* __restrict__ cz, int accum) {
  const int64_t k = item_index();
  if (k >= nsc) return;
  const int slot = leaf_slot[sc_nid[k]];
  if (slot < 0) return;
  const double den = (double)((long long)cnt[k] + (accum ? 1 : 0));
  double sum[3];
  for (int a = 0; a < 3; ++a) sum[a] = (double)(long long)vsum[3 * k + a] * (1.0 / kVelScale);
  const int br = c.res0 + 2 * apron, msk = c.res0 - 1, nv = c.sc * c.sc * c.sc;
  const int bx = sc_pos[3 * k], by = sc_pos[3 * k + 1], bz = sc_pos[3 * k + 2];
  float* ch[3] = {cx, cy, cz};
  for (int t = threadIdx.x; t < nv; t += blockDim.x) {
    const int tx = t % c.sc, ty = (t / c.sc) % c.sc, tz = t / (c.sc * c.sc);
    const int64_t vi = (((int64_t)slot * br + (((bz + tz) & msk) + apron)) * br + (((by + ty) & msk) + apron)) * br +
                       (((bx + tx) & msk) + apron);
    for (int a = 0; a < 3; ++a) {
      double v = sum[a];
      if (accum) v += (double)ch[a][vi];
      ch[a][vi] = den > 0 ? (float)(v / den) : 0.0f;
    }
  }
}

void velocity_subcell(SubcellCfg c, int64_t nsc, const int* sc_pos, const int* sc_nid, const int* cnt,
                      const unsigned long long* vsum, const int* leaf_slot, int apron, float* cx, float* cy,
                      float* cz, int accum, hipStream_t s) {
  if (nsc <= 0) return;
  const int nv = c.sc * c.sc * c.sc;
  hipLaunchKernelGGL(k_velocity_subcell, item_grid(nsc), dim3(nv < 64 ? 64 : nv), 0, s, c, nsc, sc_pos, sc_nid, cnt,
                     vsum, leaf_slot, apron, cx, cy, cz, accum);
}

// ------------------------------------------------------------------ apron
// UpdateApron<T> (O.cuh:72-126): each apron voxel of each active brick copies the interior
// voxel of the neighbour brick at the same world position, or boundval.  Neighbour slots
// come from a table built once per topology (the reference resolves them per voxel by a
// full tree descent); one launch covers several channels.
__global__ void k_neighbor_table(DevTree T, int64_t nslots, const int4* __restrict__ slot_pos,
                                 const int* __restrict__ leaf_slot, int res0, int* __restrict__ nbr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nslots * 27) return;
  const int64_t s = i / 27;
  const int k = (int)(i % 27);
  const int4 bp = slot_pos[s];
  const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
  nbr[i] = find_slot(T, bp.x + dx * res0, bp.y + dy * res0, bp.z + dz * res0, leaf_slot);
}

void neighbor_table(const DevTree& T, int64_t nslots, const int4* slot_pos, const int* leaf_slot, int res0, int* nbr,
                    hipStream_t s) {
  if (nslots <= 0) return;
  hipLaunchKernelGGL(k_neighbor_table, dim3(nblocks(nslots * 27, 256)), dim3(256), 0, s, T, nslots, slot_pos,
                     leaf_slot, res0, nbr);
}

__global__ void __launch_bounds__(256) k_apron_multi(int64_t nslots, const int* __restrict__ nbr, int res0, int apron,
                                                    ApronChannels ch) {
  __shared__ int nb[27];
  const int64_t s = item_index();
  if (s >= nslots) return;
  if (threadIdx.x < 27) nb[threadIdx.x] = nbr[s * 27 + threadIdx.x];
  __syncthreads();
  const int br = res0 + 2 * apron;
  const int nvox = br * br * br;
  const int64_t bv = (int64_t)br * br * br;
  for (int t = threadIdx.x; t < nvox; t += blockDim.x) {
    const int x = t % br, y = (t / br) % br, z = t / (br * br);
    const int gx = x - apron, gy = y - apron, gz = z - apron;
    const int ox = gx < 0 ? -1 : (gx >= res0 ? 1 : 0), oy = gy < 0 ? -1 : (gy >= res0 ? 1 : 0),
              oz = gz < 0 ? -1 : (gz >= res0 ? 1 : 0);
    if ((ox | oy | oz) == 0) continue;
    const int ns = nb[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)];
    const int64_t src = ns < 0 ? -1 : (int64_t)ns * bv +
                                          (((gz - oz * res0 + apron) * br) + (gy - oy * res0 + apron)) * br +
                                          (gx - ox * res0 + apron);
    const int64_t dst = s * bv + t;
    for (int c = 0; c < ch.n; ++c) {
      float* p = ch.ptr[c];
      p[dst] = src < 0 ? ch.bound[c] : p[src];
    }
  }
}

void apron_multi(int64_t nslots, const int* nbr, int res0, int apron, const ApronChannels& ch, hipStream_t s) {
  if (nslots <= 0 || apron <= 0 || ch.n <= 0) return;
  hipLaunchKernelGGL(k_apron_multi, item_grid(nslots), dim3(256), 0, s, nslots, nbr, res0, apron, ch);
}

}  // namespace gvdbx
