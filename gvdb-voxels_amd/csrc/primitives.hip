// primitives.hip -- wave64 scan and LSD radix sort for gfx950.
//
// Replace the reference's three-level Blelloch scan (prefixSum/prefixFixup, P.cuh:856-902,
// driven by V.cpp:4760-4819) and the Thrust uint64 sort (gvdbDeviceRadixSort,
// src/gvdb_cutils.cu:11-16).  Both are spin-free multi-kernel designs: no inter-block waits.
#include "primitives.hpp"

namespace gvdbx {

// ------------------------------------------------------------------ fills

template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

static int grid_for(int64_t n, int block, int cap = 65536) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

void fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<uint32_t>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}
void fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<uint64_t>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}
__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void copy_u32(const uint32_t* src, uint32_t* dst, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy_u32, dim3(grid_for(n, 256)), dim3(256), 0, s, src, dst, n);
}

void fill_f32(float* p, int64_t n, float v, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, p, n, v);
}

// ------------------------------------------------------------------ wide sum

__global__ void __launch_bounds__(256) k_sum_wide(const int* __restrict__ in, int64_t n,
                                                 unsigned long long* __restrict__ total) {
  __shared__ unsigned long long ws[4];
  unsigned long long acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += (unsigned int)in[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(total, ws[0] + ws[1] + ws[2] + ws[3]);
}

void sum_wide(const int* in, int64_t n, unsigned long long* total, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sum_wide, dim3(grid_for(n, 256, 1024)), dim3(256), 0, s, in, n, total);
}

// ------------------------------------------------------------------ scan


// Tile = 4 rounds x 256 threads x int4 = 4096 ints.
__global__ void __launch_bounds__(256) k_scan_reduce(const int* __restrict__ in, int64_t n, int* __restrict__ tile_sums) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int s = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t i = base + (int64_t)r * 1024 + threadIdx.x * 4;
    if (i + 3 < n) {
      int4 v = *(const int4*)(in + i);
      s += v.x + v.y + v.z + v.w;
    } else {
      for (int k = 0; k < 4; ++k)
        if (i + k < n) s += in[i + k];
    }
  }
  s = wave_sum(s);
  __shared__ int ws[4];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// Single block: exclusive scan of tile sums in place; grand total to *total.
__global__ void __launch_bounds__(256) k_scan_tiles(int* __restrict__ sums, int64_t ntiles, int* __restrict__ total) {
  int carry = 0;
  for (int64_t base = 0; base < ntiles; base += 1024) {
    int v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int64_t i = base + threadIdx.x * 4 + k;
      v[k] = i < ntiles ? sums[i] : 0;
    }
    int t = v[0] + v[1] + v[2] + v[3];
    int tot;
    int pre = block_exclusive_scan_256(t, &tot) + carry;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      int64_t i = base + threadIdx.x * 4 + k;
      if (i < ntiles) sums[i] = pre;
      pre += v[k];
    }
    carry += tot;
  }
  if (threadIdx.x == 0 && total) *total = carry;
}

__global__ void __launch_bounds__(256) k_scan_down(const int* __restrict__ in, int* __restrict__ out, int64_t n,
                                                   const int* __restrict__ tile_prefix) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  int carry = tile_prefix[blockIdx.x];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int64_t i = base + (int64_t)r * 1024 + threadIdx.x * 4;
    int v[4];
    if (i + 3 < n) {
      int4 q = *(const int4*)(in + i);
      v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
      for (int k = 0; k < 4; ++k) v[k] = (i + k < n) ? in[i + k] : 0;
    }
    int tot;
    int pre = block_exclusive_scan_256(v[0] + v[1] + v[2] + v[3], &tot) + carry;
    int o[4];
    o[0] = pre; o[1] = pre + v[0]; o[2] = o[1] + v[1]; o[3] = o[2] + v[2];
    if (i + 3 < n) {
      *(int4*)(out + i) = make_int4(o[0], o[1], o[2], o[3]);
    } else {
      for (int k = 0; k < 4; ++k)
        if (i + k < n) out[i + k] = o[k];
    }
    carry += tot;
  }
}

// Small arrays (n <= kScanSmall): one 1024-thread workgroup walks 4096-int tiles with a running
// carry -- one launch instead of three.
constexpr int64_t kScanSmall = 65536;

// n <= kScanSmall, in batches of 8 chunks of 4096 values: a batch's chunks are loaded at once
// (one memory round trip per batch), then scanned in turn
constexpr int kScanBatch = 8;

__global__ void __launch_bounds__(1024) k_scan_small(const int* __restrict__ in, int* __restrict__ out, int64_t n,
                                                     int* __restrict__ total) {
  __shared__ int ws[16];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nchunks = (int)((n + 4095) / 4096);
  int carry = 0;
  for (int b0 = 0; b0 < nchunks; b0 += kScanBatch) {
    int4 q[kScanBatch];
#pragma unroll
    for (int k = 0; k < kScanBatch; ++k) {
      const int64_t i = (int64_t)(b0 + k) * 4096 + (int64_t)tid * 4;
      q[k] = make_int4(0, 0, 0, 0);
      if (b0 + k < nchunks) {
        if (i + 3 < n) {
          q[k] = *(const int4*)(in + i);
        } else {
          q[k].x = i < n ? in[i] : 0;
          q[k].y = i + 1 < n ? in[i + 1] : 0;
          q[k].z = i + 2 < n ? in[i + 2] : 0;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kScanBatch; ++k) {
      if (b0 + k >= nchunks) continue;  // uniform over the workgroup
      const int64_t i = (int64_t)(b0 + k) * 4096 + (int64_t)tid * 4;
      const int v[4] = {q[k].x, q[k].y, q[k].z, q[k].w};
      const int t = v[0] + v[1] + v[2] + v[3];
      int x = t;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
      }
      if (lane == 63) ws[wave] = x;
      __syncthreads();
      int pre = carry, tot = 0;
#pragma unroll
      for (int w = 0; w < 16; ++w) {
        const int sw = ws[w];
        pre += (w < wave) ? sw : 0;
        tot += sw;
      }
      __syncthreads();
      pre += x - t;
      int o[4];
      o[0] = pre; o[1] = pre + v[0]; o[2] = o[1] + v[1]; o[3] = o[2] + v[2];
      if (i + 3 < n) {
        *(int4*)(out + i) = make_int4(o[0], o[1], o[2], o[3]);
      } else {
        for (int e = 0; e < 4; ++e)
          if (i + e < n) out[i + e] = o[e];
      }
      carry += tot;
    }
  }
  if (tid == 0 && total) *total = carry;
}

int64_t scan_tile_count(int64_t n) { return (n + kScanTile - 1) / kScanTile; }

void scan_tile_sums(int* sums, int64_t ntiles, int* total, hipStream_t s) {
  if (ntiles <= 0) return;
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(256), 0, s, sums, ntiles, total);
}

void exclusive_scan(const int* in, int* out, int64_t n, int* tile_buf, int* total, hipStream_t s) {
  if (n <= 0) {
    if (total) fill_u32((uint32_t*)total, 1, 0u, s);
    return;
  }
  if (n <= kScanSmall) {
    hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, in, out, n, total);
    return;
  }
  const int64_t nt = scan_tile_count(n);
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nt), dim3(256), 0, s, in, n, tile_buf);
  hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(256), 0, s, tile_buf, nt, total);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nt), dim3(256), 0, s, in, out, n, tile_buf);
}

// ------------------------------------------------------------------ radix sort

// Lanes of the wave whose digit (nbits <= 8 bits) equals this lane's (valid lanes only), from
// one ballot per bit.
__device__ __forceinline__ uint64_t match_digit(int d, bool valid, int nbits) {
  uint64_t peers = __ballot(valid);
  for (int b = 0; b < nbits; ++b) {
    const bool bit = (d >> b) & 1;
    const uint64_t bb = __ballot(bit);
    peers &= bit ? bb : ~bb;
  }
  return peers;
}

template <typename K>
__global__ void __launch_bounds__(256) k_radix_up(const K* __restrict__ keys, int64_t n, int shift, int nbits,
                                                 int* __restrict__ counts, int ntiles) {
  __shared__ int hist[256];
  hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * kRadixTile;
  const K mask = ((K)1 << nbits) - 1;
#pragma unroll 4
  for (int r = 0; r < kRadixTile / 256; ++r) {
    const int64_t i = base + r * 256 + threadIdx.x;
    if (i < n) atomicAdd(&hist[(int)((keys[i] >> shift) & mask)], 1);
  }
  __syncthreads();
  counts[(int64_t)threadIdx.x * ntiles + blockIdx.x] = hist[threadIdx.x];
}

// Per digit (one workgroup each): exclusive scan of its ntiles tile counts in place, and the
// digit's total into totals[digit].
__global__ void __launch_bounds__(256) k_digit_scan(int* __restrict__ counts, int ntiles, int* __restrict__ totals) {
  int* row = counts + (int64_t)blockIdx.x * ntiles;
  int carry = 0;
  for (int base = 0; base < ntiles; base += 256) {
    const int i = base + (int)threadIdx.x;
    const int v = i < ntiles ? row[i] : 0;
    int tot;
    const int pre = block_exclusive_scan_256(v, &tot);
    if (i < ntiles) row[i] = carry + pre;
    carry += tot;
  }
  if (threadIdx.x == 0) totals[blockIdx.x] = carry;
}

// Scatter of one tile.  Wave w ranks keys [w * 1024, (w + 1) * 1024) of the tile, 64 per round,
// against its own digit counters in LDS (the lanes of a round that share a digit take
// consecutive places, found with ballots), so the ranking needs no barrier; a per-digit scan over
// the four waves then gives every key its stable place in the tile, the keys are placed in LDS
// in digit order and written out in digit runs (consecutive threads store consecutive addresses).
// A digit's start in the output is the exclusive scan of the digit totals (formed here, one
// value per thread) plus its start among the tiles (k_digit_scan); its count in this tile is the
// next tile's start minus this one (the total for the last tile).
template <typename K, bool kHasVal>
__global__ void __launch_bounds__(256) k_radix_down(const K* __restrict__ kin, K* __restrict__ kout,
                                                   const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout,
                                                   int64_t n, int shift, int nbits, const int* __restrict__ offsets,
                                                   int ntiles, const int* __restrict__ totals) {
  constexpr int kPer = kRadixTile / 256;  // keys per lane
  const K mask = ((K)1 << nbits) - 1;
  __shared__ K sk[kRadixTile];
  __shared__ uint32_t sv[kHasVal ? kRadixTile : 1];
  __shared__ int base_of[256], lstart[256];
  __shared__ int wcnt[4][256];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t tile = blockIdx.x;
  const int64_t tbase = tile * kRadixTile;
  const int64_t wbase = tbase + (int64_t)wave * (kRadixTile / 4);
  K k[kPer];
  uint32_t v[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {  // all loads in flight before the first use
    const int64_t i = wbase + r * 64 + lane;
    k[r] = i < n ? kin[i] : (K)0;
    if (kHasVal) v[r] = i < n ? (vin ? vin[i] : (uint32_t)i) : 0u;  // no input values: the positions
  }
  const int within = offsets[(int64_t)tid * ntiles + tile];
  const int dtot = totals[tid];
  const int next = tile + 1 < ntiles ? offsets[(int64_t)tid * ntiles + tile + 1] : dtot;
#pragma unroll
  for (int w = 0; w < 4; ++w) wcnt[w][tid] = 0;
  int total;
  const int dbase = block_exclusive_scan_256(dtot, &total);
  const int ls = block_exclusive_scan_256(next - within, &total);
  base_of[tid] = dbase + within;
  lstart[tid] = ls;
  const uint64_t lt = lanemask_lt();
  int place[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    const bool valid = wbase + r * 64 + lane < n;
    const int d = (int)((k[r] >> shift) & mask);
    const uint64_t peers = match_digit(d, valid, nbits);
    const int rank = __popcll(peers & lt);
    const int c = valid ? wcnt[wave][d] : 0;
    place[r] = c + rank;
    // this wave's counters only: LDS accesses of one wave complete in program order
    if (valid && rank == 0) wcnt[wave][d] = c + __popcll(peers);
  }
  __syncthreads();
  {  // per digit: the waves' counts -> exclusive prefix over the waves, in place
    int acc = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const int c = wcnt[w][tid];
      wcnt[w][tid] = acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kPer; ++r) {
    if (wbase + r * 64 + lane >= n) continue;
    const int d = (int)((k[r] >> shift) & mask);
    const int pos = lstart[d] + wcnt[wave][d] + place[r];
    sk[pos] = k[r];
    if (kHasVal) sv[pos] = v[r];
  }
  __syncthreads();
  const int tn = (int)(n - tbase < kRadixTile ? n - tbase : kRadixTile);
  for (int j = tid; j < tn; j += 256) {
    const K key = sk[j];
    const int d = (int)((key >> shift) & mask);
    const int64_t o = (int64_t)base_of[d] + (j - lstart[d]);
    kout[o] = key;
    if (kHasVal) vout[o] = sv[j];
  }
}

// the digit-major tile counts, then the 256 digit totals
int64_t radix_count_ints(int64_t n) { return 256 * ((n + kRadixTile - 1) / kRadixTile) + 256; }

int radix_first_digit_bits(int bits) { return bits - 8 * ((bits + 7) / 8 - 1); }

// positions: vals is not read (the first pass generates the input positions, vals_alt receives
// them); first_counted: the first pass's tile counts are in counts already
template <typename K>
static int radix_sort_impl(K* keys, K* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n, int begin_bit,
                           int end_bit, int* counts, int* tile_buf, hipStream_t s, bool positions = false,
                           bool first_counted = false) {
  if (n <= 1 && !positions) return 0;
  const int ntiles = (int)((n + kRadixTile - 1) / kRadixTile);
  const int64_t ncnt = 256LL * ntiles;
  int parity = 0;
  K* kin = keys;
  K* kout = keys_alt;
  uint32_t* vin = vals;
  uint32_t* vout = vals_alt;
  int* totals = counts + ncnt;
  // digits of at most 8 bits; the first pass takes the remainder, so that 15 bits sort in 2
  // passes (7 + 8) and 18 bits in 3 (2 + 8 + 8)
  const int bits = end_bit - begin_bit;
  const int passes = (bits + 7) / 8;
  for (int pass = 0, shift = begin_bit; pass < passes; ++pass) {
    const int nbits = pass == 0 ? radix_first_digit_bits(bits) : 8;
    if (!(pass == 0 && first_counted))
      hipLaunchKernelGGL(k_radix_up<K>, dim3(ntiles), dim3(256), 0, s, kin, n, shift, nbits, counts, ntiles);
    hipLaunchKernelGGL(k_digit_scan, dim3(256), dim3(256), 0, s, counts, ntiles, totals);
    if (vals)
      hipLaunchKernelGGL((k_radix_down<K, true>), dim3(ntiles), dim3(256), 0, s, kin, kout,
                         (positions && pass == 0) ? (const uint32_t*)nullptr : vin, vout, n, shift, nbits, counts,
                         ntiles, (const int*)totals);
    else
      hipLaunchKernelGGL((k_radix_down<K, false>), dim3(ntiles), dim3(256), 0, s, kin, kout, nullptr, nullptr, n, shift,
                         nbits, counts, ntiles, (const int*)totals);
    shift += nbits;
    K* t = kin; kin = kout; kout = t;
    uint32_t* tv = vin; vin = vout; vout = tv;
    parity ^= 1;
  }
  return parity;
}

int radix_sort_u32(uint32_t* keys, uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n, int begin_bit,
                   int end_bit, int* counts, int* tile_buf, hipStream_t s) {
  return radix_sort_impl<uint32_t>(keys, keys_alt, vals, vals_alt, n, begin_bit, end_bit, counts, tile_buf, s);
}
int radix_sort_u64(uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n, int begin_bit,
                   int end_bit, int* counts, int* tile_buf, hipStream_t s) {
  return radix_sort_impl<uint64_t>(keys, keys_alt, vals, vals_alt, n, begin_bit, end_bit, counts, tile_buf, s);
}
int radix_sort_u32_positions(uint32_t* keys, uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n,
                             int end_bit, int* counts, hipStream_t s) {
  return radix_sort_impl<uint32_t>(keys, keys_alt, vals, vals_alt, n, 0, end_bit, counts, nullptr, s, true, true);
}

}  // namespace gvdbx
