// splat.hip -- trilinear mass splat into the brick atlas and the adjoints of splat and sampling.
//
// The splat is the transpose of the trilinear sampler (getVolSampleF, R.cuh:116-125, voxel v
// centred at v + 0.5): a point q with mass m adds m * w_g(q) to the 8 voxels g of its dual cell
// c = floor(q - 0.5), for every g inside an active brick; points whose own voxel floor(q) lies
// in no brick are skipped, as the sampler returns 0 there.
//
// Forward (deterministic, no device-scope atomics on the field):
//   1. k_splat_keys: key = slot(home brick) * 27 + region, where region records per axis whether
//      the point's dual cell pokes into the lower (0) or upper (2) neighbour brick or neither (1);
//      points in no brick get the last key; the launch also counts the first radix digit;
//   2. LSD radix sort of the keys with the point positions as values (the permutation sorted
//      position -> point, kept for the backward), then start[] of every key from the boundaries
//      of the sorted keys;
//   3. k_splat_place: (q, m) of each point as float4 at its sorted position (the inverse
//      permutation scatters them);
//   4. k_splat_brick: one wave per brick pulls its own points plus exactly those points of the
//      26 neighbours whose region reaches it, and sums m * w in LDS as 64-bit integers with a
//      per-brick power-of-two scale taken from max |m| -- integer sums are order independent, so
//      the FP32 result is bit-reproducible whatever the order inside a key; the interior is
//      written once.
// Backward of the splat (gather): dL/dm_i = sum_g w_g(q_i) G(g), dL/dq_i = m_i sum_g grad w_g G(g),
// one thread per point, visiting the points in the brick order of the last forward splat of the
// same buffer when there is one (any permutation gives the same result; brick order keeps
// the reads of G local).
// Backward of sampling: dL/dq_i = dout_i * grad sample(q_i) (same padded-brick filter as the
// sampler); dL/dfield = splat of dout (the forward above).
#include "splat.hpp"

#include "primitives.hpp"

namespace gvdbx {

namespace {

inline int nblocks(int64_t n, int bs) { return (int)((n + bs - 1) / bs > 0 ? (n + bs - 1) / bs : 1); }

__device__ __forceinline__ int region_of(int l, float frac, int res) {
  return (l == 0 && frac < 0.5f) ? 0 : ((l == res - 1 && frac >= 0.5f) ? 2 : 1);
}

}  // namespace

__device__ __forceinline__ uint32_t splat_key(const DevTree& T, float px, float py, float pz,
                                              const int* __restrict__ leaf_slot, int lgres, uint32_t sentinel) {
  if (!(px >= 0.f && py >= 0.f && pz >= 0.f)) return sentinel;
  const int ix = (int)px, iy = (int)py, iz = (int)pz;
  const int slot = find_slot(T, ix, iy, iz, leaf_slot);
  if (slot < 0) return sentinel;
  const int res = 1 << lgres, m = res - 1;
  const int rx = region_of(ix & m, px - (float)ix, res), ry = region_of(iy & m, py - (float)iy, res),
            rz = region_of(iz & m, pz - (float)iz, res);
  return (uint32_t)slot * 27u + (uint32_t)(rz * 9 + ry * 3 + rx);
}

// Key of every point, and per tile of kRadixTile points the counts of the first radix digit
// (the low nbits bits), digit-major into counts[digit * ntiles + tile], as the radix sort's
// first pass would count them.
__global__ void __launch_bounds__(kKeyBlock) k_splat_keys(DevTree T, const float* __restrict__ q, int64_t n,
                                                    const int* __restrict__ leaf_slot, int lgres, uint32_t sentinel,
                                                    uint32_t* __restrict__ keys, int nbits, int* __restrict__ counts,
                                                    int ntiles) {
  __shared__ int hist[256];
  if (threadIdx.x < 256) hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  const uint32_t mask = (1u << nbits) - 1u;
#pragma unroll
  for (int r = 0; r < kRadixTile / kKeyBlock; ++r) {
    const int64_t i = base + r * kKeyBlock + threadIdx.x;
    if (i >= n) continue;
    const uint32_t key = splat_key(T, q[3 * i], q[3 * i + 1], q[3 * i + 2], leaf_slot, lgres, sentinel);
    keys[i] = key;
    atomicAdd(&hist[key & mask], 1);
  }
  __syncthreads();
  if (threadIdx.x < 256) counts[(int64_t)threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

// inv[perm[j]] = j: the sorted position of every point
__global__ void __launch_bounds__(256) k_splat_inverse(const uint32_t* __restrict__ perm, int64_t n,
                                                       uint32_t* __restrict__ inv) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) inv[perm[j]] = (uint32_t)j;
}

// payload[inv[i]] = (q, m) of point i (m = 1 without masses): the points are read in their own
// order and scattered, since a wave's scattered writes do not wait on memory as its scattered
// reads would
__global__ void __launch_bounds__(256) k_splat_place(const float* __restrict__ q, const float* __restrict__ mass,
                                                     int64_t n, const uint32_t* __restrict__ inv,
                                                     float4* __restrict__ payload) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  payload[inv[i]] = make_float4(q[3 * i], q[3 * i + 1], q[3 * i + 2], mass ? mass[i] : 1.f);
}

// Per axis, the 5 (neighbour offset, region) pairs whose points reach this brick.
__device__ __forceinline__ void axis_combo(int k, int& d, int& r) {
  d = k < 3 ? 0 : (k == 3 ? -1 : 1);
  r = k < 3 ? k : (k == 3 ? 2 : 0);
}

template <int kLg>
__global__ void __launch_bounds__(64) k_splat_brick(int64_t nslots, const int* __restrict__ start,
                                                   const float4* __restrict__ payload,
                                                   const int* __restrict__ nbr, const int4* __restrict__ slot_pos,
                                                   int apron, float* __restrict__ chan, int accum) {
  constexpr int kRes = 1 << kLg, kVox = kRes * kRes * kRes, kRanges = 125;
  __shared__ long long acc[kVox];
  __shared__ int rbeg[kRanges];      // non-empty (direction, region) ranges: first sorted position
  __shared__ int rpre[kRanges + 1];  // their starts in the brick's virtual point list
  const int64_t s = item_index();  // one wave per brick
  if (s >= nslots) return;
  const int lane = threadIdx.x;
  for (int v = lane; v < kVox; v += 64) acc[v] = 0;
  const int nbv = lane < 27 ? nbr[(int64_t)s * 27 + lane] : -1;
  // lane l owns ranges l and l + 64; empty ones are dropped by ballot compaction
  int beg[2] = {0, 0}, len[2] = {0, 0};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int r = lane + 64 * t;
    int dx, dy, dz, rx, ry, rz;
    axis_combo(r % 5, dx, rx);
    axis_combo((r / 5) % 5, dy, ry);
    axis_combo((r / 25) % 5, dz, rz);
    const int nb = __shfl(nbv, (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1), 64);
    if (r < kRanges && nb >= 0) {
      const int64_t k = (int64_t)nb * 27 + rz * 9 + ry * 3 + rx;
      beg[t] = start[k];
      len[t] = start[k + 1] - beg[t];
    }
  }
  int base = 0, nr = 0;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    int incl = len[t];
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(incl, o, 64);
      if (lane >= o) incl += u;
    }
    const uint64_t m = __ballot(len[t] > 0);
    if (len[t] > 0) {
      const int e = nr + __popcll(m & lanemask_lt());
      rbeg[e] = beg[t];
      rpre[e] = base + incl - len[t];
    }
    nr += __popcll(m);
    base += __shfl(incl, 63, 64);
  }
  const int total = base;
  if (lane == 0) rpre[nr] = total;
  __syncthreads();
  const int br = kRes + 2 * apron;
  float* dst = chan + (int64_t)s * br * br * br;
  if (total == 0) {  // no point reaches this brick
    if (!accum)
      for (int v = lane; v < kVox; v += 64) {
        const int x = v & (kRes - 1), y = (v >> kLg) & (kRes - 1), z = v >> (2 * kLg);
        dst[((z + apron) * br + (y + apron)) * br + (x + apron)] = 0.f;
      }
    return;
  }
  auto point_of = [&](int j) -> float4 {
    int lo = 0, hi = nr - 1;  // last range starting at or before j
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rpre[mid] <= j) lo = mid;
      else hi = mid - 1;
    }
    return payload[rbeg[lo] + (j - rpre[lo])];
  };
  // per-brick fixed-point scale: each voxel receives at most `total` terms of |m w| <= max|m|;
  // the lane's first point stays in registers between the two passes
  const float4 p0 = lane < total ? point_of(lane) : make_float4(0.f, 0.f, 0.f, 0.f);
  unsigned int mb = __float_as_uint(fabsf(p0.w));
  for (int j = lane + 64; j < total; j += 64) {
    const unsigned int b = __float_as_uint(fabsf(point_of(j).w));
    mb = b > mb ? b : mb;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned int u = __shfl_xor(mb, o, 64);
    mb = u > mb ? u : mb;
  }
  const double bound = (double)__uint_as_float(mb) * (double)total;
  int e = 0;
  if (bound > 0.0) (void)frexp(bound, &e);  // bound < 2^e
  const int S = 62 - e;                     // every partial sum stays below 2^62
  const int4 bp = slot_pos[s];
  for (int j = lane; j < total; j += 64) {
    const float4 p = j == lane ? p0 : point_of(j);
    const double x = (double)p.x - 0.5, y = (double)p.y - 0.5, z = (double)p.z - 0.5;
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const int cx = (int)fx - bp.x, cy = (int)fy - bp.y, cz = (int)fz - bp.z;
    const double ax = x - fx, ay = y - fy, az = z - fz;
    const double ms = ldexp((double)p.w, S);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
      const int gx = cx + dx, gy = cy + dy, gz = cz + dz;
      if ((unsigned)gx >= (unsigned)kRes || (unsigned)gy >= (unsigned)kRes || (unsigned)gz >= (unsigned)kRes) continue;
      const double w = (dx ? ax : 1.0 - ax) * (dy ? ay : 1.0 - ay) * (dz ? az : 1.0 - az);
      const long long t = __double2ll_rn(ms * w);
      if (t != 0) atomicAdd((unsigned long long*)&acc[(gz * kRes + gy) * kRes + gx], (unsigned long long)t);
    }
  }
  __syncthreads();
  for (int v = lane; v < kVox; v += 64) {
    const int x = v & (kRes - 1), y = (v >> kLg) & (kRes - 1), z = v >> (2 * kLg);
    float* o = dst + ((z + apron) * br + (y + apron)) * br + (x + apron);
    const double val = ldexp((double)acc[v], -S);
    *o = accum ? (float)((double)*o + val) : (float)val;
  }
}

// dL/dm and dL/dq of the splat from G = dL/drho (interior voxels of chan).
__global__ void __launch_bounds__(256) k_splat_backward(DevTree T, const int* __restrict__ leaf_slot,
                                                       const int* __restrict__ nbr, int lgres, int apron,
                                                       const float* __restrict__ chan, const float* __restrict__ q,
                                                       const float* __restrict__ mass, int64_t n,
                                                       const uint32_t* __restrict__ order, float* __restrict__ dq,
                                                       float* __restrict__ dm) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int64_t i = order ? (int64_t)order[j] : j;  // a permutation of [0, n): brick order for locality
  const float px = q[3 * i], py = q[3 * i + 1], pz = q[3 * i + 2];
  double s = 0, gx = 0, gy = 0, gz = 0;
  const double m = mass ? (double)mass[i] : 1.0;
  if (px >= 0.f && py >= 0.f && pz >= 0.f) {
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    const int slot = find_slot(T, ix, iy, iz, leaf_slot);
    if (slot >= 0) {
      const int res = 1 << lgres, msk = ~(res - 1), br = res + 2 * apron;
      const int64_t bv = (int64_t)br * br * br;
      const double x = (double)px - 0.5, y = (double)py - 0.5, z = (double)pz - 0.5;
      const double fx = floor(x), fy = floor(y), fz = floor(z);
      const int cx = (int)fx - (ix & msk), cy = (int)fy - (iy & msk), cz = (int)fz - (iz & msk);
      const double ax = x - fx, ay = y - fy, az = z - fz;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        const int lx = cx + dx, ly = cy + dy, lz = cz + dz;  // in [-1, res]
        const int ox = lx < 0 ? -1 : (lx >= res ? 1 : 0), oy = ly < 0 ? -1 : (ly >= res ? 1 : 0),
                  oz = lz < 0 ? -1 : (lz >= res ? 1 : 0);
        const int nb = (ox | oy | oz) == 0 ? slot : nbr[(int64_t)slot * 27 + (oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)];
        if (nb < 0) continue;
        const double G = (double)chan[nb * bv + (((lz - oz * res + apron) * br) + (ly - oy * res + apron)) * br +
                                      (lx - ox * res + apron)];
        const double wx = dx ? ax : 1.0 - ax, wy = dy ? ay : 1.0 - ay, wz = dz ? az : 1.0 - az;
        s += wx * wy * wz * G;
        gx += (dx ? 1.0 : -1.0) * wy * wz * G;
        gy += wx * (dy ? 1.0 : -1.0) * wz * G;
        gz += wx * wy * (dz ? 1.0 : -1.0) * G;
      }
    }
  }
  if (dm) dm[i] = (float)s;
  if (dq) {
    dq[3 * i] = (float)(m * gx);
    dq[3 * i + 1] = (float)(m * gy);
    dq[3 * i + 2] = (float)(m * gz);
  }
}

// dq_i = dout_i * grad sample(q_i), the sampler's own padded-brick filter (apron included).
__global__ void __launch_bounds__(256) k_sample_backward(DevTree T, const int* __restrict__ leaf_slot, int res0,
                                                        int apron, const float* __restrict__ chan,
                                                        const float* __restrict__ q, const float* __restrict__ dout,
                                                        int64_t n, float* __restrict__ dq) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = q[3 * i], py = q[3 * i + 1], pz = q[3 * i + 2];
  double gx = 0, gy = 0, gz = 0;
  if (px >= 0.f && py >= 0.f && pz >= 0.f) {
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    const int slot = find_slot(T, ix, iy, iz, leaf_slot);
    if (slot >= 0) {
      const int msk = ~(res0 - 1), br = res0 + 2 * apron;
      const double x = (double)px - (ix & msk) + apron - 0.5, y = (double)py - (iy & msk) + apron - 0.5,
                   z = (double)pz - (iz & msk) + apron - 0.5;
      const double fx = floor(x), fy = floor(y), fz = floor(z);
      const int x0 = (int)fx, y0 = (int)fy, z0 = (int)fz;
      const double ax = x - fx, ay = y - fy, az = z - fz;
      const float* b = chan + (int64_t)slot * br * br * br;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int dx = k & 1, dy = (k >> 1) & 1, dz = k >> 2;
        const double v = (double)b[((z0 + dz) * br + (y0 + dy)) * br + (x0 + dx)];
        const double wx = dx ? ax : 1.0 - ax, wy = dy ? ay : 1.0 - ay, wz = dz ? az : 1.0 - az;
        gx += (dx ? 1.0 : -1.0) * wy * wz * v;
        gy += wx * (dy ? 1.0 : -1.0) * wz * v;
        gz += wx * wy * (dz ? 1.0 : -1.0) * v;
      }
    }
  }
  const double d = (double)dout[i];
  dq[3 * i] = (float)(d * gx);
  dq[3 * i + 1] = (float)(d * gy);
  dq[3 * i + 2] = (float)(d * gz);
}

void splat_keys(const DevTree& T, const float* q, int64_t n, const int* leaf_slot, int lgres, uint32_t sentinel,
                uint32_t* keys, int bits, int* counts, hipStream_t s) {
  if (n <= 0) return;
  const int ntiles = (int)((n + kRadixTile - 1) / kRadixTile);
  hipLaunchKernelGGL(k_splat_keys, dim3(ntiles), dim3(kKeyBlock), 0, s, T, q, n, leaf_slot, lgres, sentinel, keys,
                     radix_first_digit_bits(bits), counts, ntiles);
}

void splat_place(const float* q, const float* mass, int64_t n, const uint32_t* perm, uint32_t* inv, float4* payload,
                 hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_splat_inverse, dim3(nblocks(n, 256)), dim3(256), 0, s, perm, n, inv);
  hipLaunchKernelGGL(k_splat_place, dim3(nblocks(n, 256)), dim3(256), 0, s, q, mass, n, (const uint32_t*)inv,
                     payload);
}

bool splat_brick_supported(int lgres) { return lgres >= 1 && lgres <= 4; }

void splat_brick(int64_t nslots, const int* start, const float4* payload, const int* nbr, const int4* slot_pos,
                 int lgres, int apron, float* chan, int accum, hipStream_t s) {
  if (nslots <= 0) return;
  const dim3 g = item_grid(nslots), b(64);
  switch (lgres) {
    case 1: hipLaunchKernelGGL(k_splat_brick<1>, g, b, 0, s, nslots, start, payload, nbr, slot_pos, apron, chan, accum); break;
    case 2: hipLaunchKernelGGL(k_splat_brick<2>, g, b, 0, s, nslots, start, payload, nbr, slot_pos, apron, chan, accum); break;
    case 3: hipLaunchKernelGGL(k_splat_brick<3>, g, b, 0, s, nslots, start, payload, nbr, slot_pos, apron, chan, accum); break;
    default: hipLaunchKernelGGL(k_splat_brick<4>, g, b, 0, s, nslots, start, payload, nbr, slot_pos, apron, chan, accum); break;
  }
}

void splat_backward(const DevTree& T, const int* leaf_slot, const int* nbr, int lgres, int apron, const float* chan,
                    const float* q, const float* mass, int64_t n, const uint32_t* order, float* dq, float* dm,
                    hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_splat_backward, dim3(nblocks(n, 256)), dim3(256), 0, s, T, leaf_slot, nbr, lgres, apron, chan,
                     q, mass, n, order, dq, dm);
}

void sample_backward(const DevTree& T, const int* leaf_slot, int res0, int apron, const float* chan, const float* q,
                     const float* dout, int64_t n, float* dq, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sample_backward, dim3(nblocks(n, 256)), dim3(256), 0, s, T, leaf_slot, res0, apron, chan, q,
                     dout, n, dq);
}

// ------------------------------------------------------------------ ScatterDensity
// gvdbScatterPointDensity (P.cuh:516-573) with the leaf of q = p + trans found as
// gvdbInsertPoints does (P.cuh:19-36): for a valid point in a leaf with corner b, p' = q - b,
// pi = trunc(p'); the voxel pi, and with expand its 6 face neighbours (unclamped: they may lie
// in the brick's apron; dropped without an apron), gains (1 - c)^3 = 1 + c(-3 + c(3 - c)),
// c = |v - p'|^2 / r^2 (P.cuh:514-521).  The reference's read-modify-write races; here every
// contribution is summed, as a 64-bit fixed-point integer (scale 2^shift) per padded voxel, so
// the sum does not depend on the order of the atomics.
__device__ __forceinline__ float scatter_weight(float px, float py, float pz, float vx, float vy, float vz, float r) {
  const float bx = vx - px, by = vy - py, bz = vz - pz;
  const float c = (bx * bx + by * by + bz * bz) / (r * r);
  return 1.0f + c * (-3.0f + c * (3.0f - c));
}

// cmode: colour part of the same kernel (P.cuh:553-572) -- 0 none; 1 avgColor: count and r, g, b
// sums per interior voxel (cbuf: 4 x u64 per voxel); 2 the point colour itself: the largest
// index + 1 of the points in the voxel (cbuf: u32 per voxel).
__global__ void __launch_bounds__(256) k_scatter_density(DevTree T, PointsView P, const int* __restrict__ leaf_slot,
                                                         int lgres, int apron, float r, int expand, double scale,
                                                         unsigned long long* __restrict__ acc, int cmode,
                                                         void* __restrict__ cbuf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P.n) return;
  const float3 p = load_f3(P.pos, i, P.pos_off, P.pos_stride);
  const float q[3] = {p.x + P.tx, p.y + P.ty, p.z + P.tz};
  // invalid (negative), and far beyond any root (NOHIT = 1e10): no leaf
  if (!(q[0] >= 0.f && q[1] >= 0.f && q[2] >= 0.f && q[0] < 1073741824.f && q[1] < 1073741824.f &&
        q[2] < 1073741824.f))
    return;
  const int iq[3] = {(int)q[0], (int)q[1], (int)q[2]};
  const int slot = find_slot(T, iq[0], iq[1], iq[2], leaf_slot);
  if (slot < 0) return;
  const int res = 1 << lgres, br = res + 2 * apron;
  float pl[3];
  int pi[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    pl[a] = q[a] - (float)(iq[a] & ~(res - 1));  // brick corner: nodes are aligned to their range
    pi[a] = (int)pl[a];
  }
  if (cmode && (unsigned)pi[0] < (unsigned)res && (unsigned)pi[1] < (unsigned)res && (unsigned)pi[2] < (unsigned)res) {
    const int64_t vx = ((int64_t)slot << (3 * lgres)) + (((pi[2] << lgres) + pi[1]) << lgres) + pi[0];
    if (cmode == 1) {
      const uint32_t w = load_u32(P.clr, i, P.clr_off, P.clr_stride);
      unsigned long long* cs = (unsigned long long*)cbuf + 4 * vx;
      atomicAdd(cs, 1ull);
      atomicAdd(cs + 1, (unsigned long long)(w & 255u));
      atomicAdd(cs + 2, (unsigned long long)((w >> 8) & 255u));
      atomicAdd(cs + 3, (unsigned long long)((w >> 16) & 255u));
    } else {
      atomicMax((uint32_t*)cbuf + vx, (uint32_t)(i + 1));
    }
  }
  unsigned long long* dst = acc + (int64_t)slot * br * br * br;
  const int nk = expand ? 7 : 1;
  for (int k = 0; k < nk; ++k) {
    const int ax = k == 0 ? -1 : (k - 1) >> 1, d = (k & 1) ? -1 : 1;
    int v[3] = {pi[0], pi[1], pi[2]};
    if (ax >= 0) v[ax] += d;
    if (v[0] < -apron || v[1] < -apron || v[2] < -apron || v[0] >= res + apron || v[1] >= res + apron ||
        v[2] >= res + apron)
      continue;
    const float w = scatter_weight(pl[0], pl[1], pl[2], (float)v[0], (float)v[1], (float)v[2], r);
    const long long fx = __double2ll_rn((double)w * scale);
    atomicAdd(&dst[((v[2] + apron) * br + (v[1] + apron)) * br + (v[0] + apron)], (unsigned long long)fx);
  }
}

__global__ void __launch_bounds__(256) k_scatter_apply(const unsigned long long* __restrict__ acc, int64_t n,
                                                       double inv_scale, float* __restrict__ chan) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const long long v = (long long)acc[i];
    if (v != 0) chan[i] += (float)((double)v * inv_scale);
  }
}

// Colour words of the voxels the points reached (gvdbScatterPointAvgCol, P.cuh:610-641, for
// cmode 1: (r, g, b sums / count, 255); cmode 2: the colour of the recorded point); voxels no
// point reached keep their word.
__global__ void __launch_bounds__(256) k_scatter_colour_apply(PointsView P, int64_t nin, int lgres, int apron, int cmode,
                                                              const void* __restrict__ cbuf,
                                                              uint32_t* __restrict__ chan) {
  const int res = 1 << lgres, br = res + 2 * apron;
  for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nin; v += (int64_t)gridDim.x * blockDim.x) {
    uint32_t word;
    if (cmode == 1) {
      const unsigned long long* cs = (const unsigned long long*)cbuf + 4 * v;
      const unsigned long long n = cs[0];
      if (n == 0) continue;
      word = (uint32_t)(cs[1] / n) | (uint32_t)(cs[2] / n) << 8 | (uint32_t)(cs[3] / n) << 16 | 255u << 24;
    } else {
      const uint32_t k = ((const uint32_t*)cbuf)[v];
      if (k == 0) continue;
      word = load_u32(P.clr, (int64_t)k - 1, P.clr_off, P.clr_stride);
    }
    const int64_t slot = v >> (3 * lgres);
    const int l = (int)(v & ((1 << (3 * lgres)) - 1));
    const int x = l & (res - 1), y = (l >> lgres) & (res - 1), z = l >> (2 * lgres);
    chan[((slot * br + z + apron) * br + y + apron) * br + x + apron] = word;
  }
}

void scatter_density(const DevTree& T, const PointsView& P, const int* leaf_slot, int lgres, int apron, float r,
                     int expand, int shift, int64_t nslots, unsigned long long* acc, float* chan, int cmode,
                     void* cbuf, uint32_t* cchan, hipStream_t s) {
  const int br = (1 << lgres) + 2 * apron;
  const int64_t nvox = nslots * br * br * br;
  if (nvox <= 0) return;
  const int64_t nin = nslots << (3 * lgres);
  fill_u64((uint64_t*)acc, nvox, 0ull, s);
  if (cmode == 1) fill_u64((uint64_t*)cbuf, 4 * nin, 0ull, s);
  if (cmode == 2) fill_u32((uint32_t*)cbuf, nin, 0u, s);
  if (P.n > 0)
    hipLaunchKernelGGL(k_scatter_density, dim3(nblocks(P.n, 256)), dim3(256), 0, s, T, P, leaf_slot, lgres, apron, r,
                       expand, ldexp(1.0, shift), acc, cmode, cbuf);
  const int64_t g = (nvox + 255) / 256;
  hipLaunchKernelGGL(k_scatter_apply, dim3((unsigned)(g < 65536 ? g : 65536)), dim3(256), 0, s, acc, nvox,
                     ldexp(1.0, -shift), chan);
  if (cmode) {
    const int64_t gc = (nin + 255) / 256;
    hipLaunchKernelGGL(k_scatter_colour_apply, dim3((unsigned)(gc < 65536 ? gc : 65536)), dim3(256), 0, s, P, nin,
                       lgres, apron, cmode, (const void*)cbuf, cchan);
  }
}

}  // namespace gvdbx
