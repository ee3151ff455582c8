// splat.hip -- trilinear mass splat into the brick atlas and the adjoints of splat and sampling.
//
// The splat is the transpose of the trilinear sampler (getVolSampleF, R.cuh:116-125, voxel v
// centred at v + 0.5): a point q with mass m adds m * w_g(q) to the 8 voxels g of its dual cell
// c = floor(q - 0.5), for every g inside an active brick; points whose own voxel floor(q) lies
// in no brick are skipped, as the sampler returns 0 there.
//
// Forward (deterministic, no device-scope atomics):
//   1. k_splat_keys: key = slot(home brick) * 27 + region, where region records per axis whether
//      the point's dual cell pokes into the lower (0) or upper (2) neighbour brick or neither (1);
//   2. stable radix sort of (key, point index) and start[] of every key;
//   3. k_splat_brick: one workgroup per brick pulls its own points plus exactly those points of
//      the 26 neighbours whose region reaches it, and sums m * w in LDS as 64-bit integers with a
//      per-brick power-of-two scale taken from max |m| -- integer sums are order independent, so
//      the FP32 result is bit-reproducible; the interior is written once, no global atomics.
// Backward of the splat (gather): dL/dm_i = sum_g w_g(q_i) G(g), dL/dq_i = m_i sum_g grad w_g G(g).
// Backward of sampling: dL/dq_i = dout_i * grad sample(q_i) (same padded-brick filter as the
// sampler); dL/dfield = splat of dout (the forward above).
#include "splat.hpp"

namespace gvdbx {

namespace {

inline int nblocks(int64_t n, int bs) { return (int)((n + bs - 1) / bs > 0 ? (n + bs - 1) / bs : 1); }

__device__ __forceinline__ int region_of(int l, float frac, int res) {
  return (l == 0 && frac < 0.5f) ? 0 : ((l == res - 1 && frac >= 0.5f) ? 2 : 1);
}

}  // namespace

__global__ void __launch_bounds__(256) k_splat_keys(DevTree T, const float* __restrict__ q, int64_t n,
                                                    const int* __restrict__ leaf_slot, int lgres, uint32_t sentinel,
                                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = q[3 * i], py = q[3 * i + 1], pz = q[3 * i + 2];
  uint32_t key = sentinel;
  if (px >= 0.f && py >= 0.f && pz >= 0.f) {
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    const int64_t lf = descend(T, ix, iy, iz, 0);
    const int slot = lf < 0 ? -1 : leaf_slot[lf];
    if (slot >= 0) {
      const int res = 1 << lgres, m = res - 1;
      const int rx = region_of(ix & m, px - (float)ix, res), ry = region_of(iy & m, py - (float)iy, res),
                rz = region_of(iz & m, pz - (float)iz, res);
      key = (uint32_t)slot * 27u + (uint32_t)(rz * 9 + ry * 3 + rx);
    }
  }
  keys[i] = key;
  vals[i] = (uint32_t)i;
}

// Per axis, the 5 (neighbour offset, region) pairs whose points reach this brick.
__device__ __forceinline__ void axis_combo(int k, int& d, int& r) {
  d = k < 3 ? 0 : (k == 3 ? -1 : 1);
  r = k < 3 ? k : (k == 3 ? 2 : 0);
}

template <int kLg>
__global__ void __launch_bounds__(256) k_splat_brick(const int* __restrict__ start, const uint32_t* __restrict__ svals,
                                                    const float* __restrict__ q, const float* __restrict__ mass,
                                                    const int* __restrict__ nbr, const int4* __restrict__ slot_pos,
                                                    int apron, float* __restrict__ chan, int accum) {
  constexpr int kRes = 1 << kLg, kVox = kRes * kRes * kRes, kRanges = 125;
  __shared__ long long acc[kVox];
  __shared__ int rbeg[kRanges];
  __shared__ int rpre[kRanges + 1];
  __shared__ unsigned int mbits;
  const int s = blockIdx.x, tid = threadIdx.x;
  for (int v = tid; v < kVox; v += blockDim.x) acc[v] = 0;
  if (tid == 0) mbits = 0u;
  // range table: 125 (direction, region) key ranges, prefix of their lengths by wave 0
  int len = 0, beg = 0;
  if (tid < kRanges) {
    int dx, dy, dz, rx, ry, rz;
    axis_combo(tid % 5, dx, rx);
    axis_combo((tid / 5) % 5, dy, ry);
    axis_combo(tid / 25, dz, rz);
    const int nb = nbr[(int64_t)s * 27 + (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)];
    if (nb >= 0) {
      const int64_t k = (int64_t)nb * 27 + rz * 9 + ry * 3 + rx;
      beg = start[k];
      len = start[k + 1] - beg;
    }
    rbeg[tid] = beg;
  }
  if (tid == 0) rpre[0] = 0;
  if (tid < kRanges) rpre[tid + 1] = len;
  __syncthreads();
  if (tid < 64) {
    // wave 0 scans the lengths, two ranges per lane: lane l holds ranges 2l and 2l + 1,
    // whose lengths sit in rpre[2l + 1] and rpre[2l + 2] (indices up to kRanges exist)
    const int ia = 2 * tid + 1, ib = 2 * tid + 2;
    const int a = ia <= kRanges ? rpre[ia] : 0, b = ib <= kRanges ? rpre[ib] : 0;
    int incl = a + b;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int t = __shfl_up(incl, o, 64);
      if (tid >= o) incl += t;
    }
    if (ia <= kRanges) rpre[ia] = incl - b;
    if (ib <= kRanges) rpre[ib] = incl;
  }
  __syncthreads();
  const int total = rpre[kRanges];
  if (total == 0) {
    // no point reaches this brick
    const int br = kRes + 2 * apron;
    float* dst = chan + (int64_t)s * br * br * br;
    if (!accum)
      for (int v = tid; v < kVox; v += blockDim.x) {
        const int x = v & (kRes - 1), y = (v >> kLg) & (kRes - 1), z = v >> (2 * kLg);
        dst[((z + apron) * br + (y + apron)) * br + (x + apron)] = 0.f;
      }
    return;
  }
  auto point_of = [&](int j) -> uint32_t {
    int lo = 0, hi = kRanges - 1;  // last range with rpre[r] <= j
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rpre[mid] <= j) lo = mid;
      else hi = mid - 1;
    }
    return svals[rbeg[lo] + (j - rpre[lo])];
  };
  // per-brick fixed-point scale: every voxel receives at most `total` terms of |m w| <= max|m|
  if (mass) {
    unsigned int mb = 0u;
    for (int j = tid; j < total; j += blockDim.x) {
      const unsigned int b = __float_as_uint(fabsf(mass[point_of(j)]));
      mb = b > mb ? b : mb;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const unsigned int t = __shfl_xor(mb, o, 64);
      mb = t > mb ? t : mb;
    }
    if ((tid & 63) == 0) atomicMax(&mbits, mb);
  }
  __syncthreads();
  const double maxm = mass ? (double)__uint_as_float(mbits) : 1.0;
  const double bound = maxm * (double)total;
  int e = 0;
  if (bound > 0.0) (void)frexp(bound, &e);  // bound < 2^e
  const int S = 62 - e;                     // every partial sum stays below 2^62
  const int4 bp = slot_pos[s];
  for (int j = tid; j < total; j += blockDim.x) {
    const uint32_t i = point_of(j);
    const double m = mass ? (double)mass[i] : 1.0;
    const double x = (double)q[3 * (int64_t)i] - 0.5, y = (double)q[3 * (int64_t)i + 1] - 0.5,
                 z = (double)q[3 * (int64_t)i + 2] - 0.5;
    const double fx = floor(x), fy = floor(y), fz = floor(z);
    const int cx = (int)fx - bp.x, cy = (int)fy - bp.y, cz = (int)fz - bp.z;
    const double ax = x - fx, ay = y - fy, az = z - fz;
    const double ms = ldexp(m, S);
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
  const int br = kRes + 2 * apron;
  float* dst = chan + (int64_t)s * br * br * br;
  for (int v = tid; v < kVox; v += blockDim.x) {
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
                                                       float* __restrict__ dq, float* __restrict__ dm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float px = q[3 * i], py = q[3 * i + 1], pz = q[3 * i + 2];
  double s = 0, gx = 0, gy = 0, gz = 0;
  const double m = mass ? (double)mass[i] : 1.0;
  if (px >= 0.f && py >= 0.f && pz >= 0.f) {
    const int ix = (int)px, iy = (int)py, iz = (int)pz;
    const int64_t lf = descend(T, ix, iy, iz, 0);
    const int slot = lf < 0 ? -1 : leaf_slot[lf];
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
    const int64_t lf = descend(T, ix, iy, iz, 0);
    const int slot = lf < 0 ? -1 : leaf_slot[lf];
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
                uint32_t* keys, uint32_t* vals, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_splat_keys, dim3(nblocks(n, 256)), dim3(256), 0, s, T, q, n, leaf_slot, lgres, sentinel, keys,
                     vals);
}

bool splat_brick_supported(int lgres) { return lgres >= 1 && lgres <= 4; }

void splat_brick(int64_t nslots, const int* start, const uint32_t* svals, const float* q, const float* mass,
                 const int* nbr, const int4* slot_pos, int lgres, int apron, float* chan, int accum, hipStream_t s) {
  if (nslots <= 0) return;
  const dim3 g((unsigned)nslots), b(256);
  switch (lgres) {
    case 1: hipLaunchKernelGGL(k_splat_brick<1>, g, b, 0, s, start, svals, q, mass, nbr, slot_pos, apron, chan, accum); break;
    case 2: hipLaunchKernelGGL(k_splat_brick<2>, g, b, 0, s, start, svals, q, mass, nbr, slot_pos, apron, chan, accum); break;
    case 3: hipLaunchKernelGGL(k_splat_brick<3>, g, b, 0, s, start, svals, q, mass, nbr, slot_pos, apron, chan, accum); break;
    default: hipLaunchKernelGGL(k_splat_brick<4>, g, b, 0, s, start, svals, q, mass, nbr, slot_pos, apron, chan, accum); break;
  }
}

void splat_backward(const DevTree& T, const int* leaf_slot, const int* nbr, int lgres, int apron, const float* chan,
                    const float* q, const float* mass, int64_t n, float* dq, float* dm, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_splat_backward, dim3(nblocks(n, 256)), dim3(256), 0, s, T, leaf_slot, nbr, lgres, apron, chan,
                     q, mass, n, dq, dm);
}

void sample_backward(const DevTree& T, const int* leaf_slot, int res0, int apron, const float* chan, const float* q,
                     const float* dout, int64_t n, float* dq, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sample_backward, dim3(nblocks(n, 256)), dim3(256), 0, s, T, leaf_slot, res0, apron, chan, q,
                     dout, n, dq);
}

}  // namespace gvdbx
