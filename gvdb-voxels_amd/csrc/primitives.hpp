// primitives.hpp -- scan / radix sort / fill launchers (device code in primitives.hip).
#pragma once
#include "common.hpp"

namespace gvdbx {

constexpr int kScanTile = 4096;
constexpr int kRadixTile = 4096;
constexpr int kKeyBlock = 1024;  // threads of the kernels that make sort keys with the first digit's tile counts

void fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
void fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t s);
void fill_f32(float* p, int64_t n, float v, hipStream_t s);
// device-to-device copy by a kernel (records into a graph as a kernel node)
void copy_u32(const uint32_t* src, uint32_t* dst, int64_t n, hipStream_t s);

// Exclusive scan of n int32 values (in may equal out).  tile_buf holds scan_tile_count(n)
// ints; *total (device) receives the sum when non-null.
int64_t scan_tile_count(int64_t n);
void exclusive_scan(const int* in, int* out, int64_t n, int* tile_buf, int* total, hipStream_t s);
// Exclusive scan of ntiles block sums in place (one workgroup); their total to *total.
void scan_tile_sums(int* sums, int64_t ntiles, int* total, hipStream_t s);

// *total (device, zeroed by the caller) += sum of n non-negative int32 values, in 64 bits: the
// scan's 32-bit total wraps once the sum passes 2^32.
void sum_wide(const int* in, int64_t n, unsigned long long* total, hipStream_t s);

// Stable LSD radix sort over bits [begin_bit, end_bit) in 8-bit digits, ping-ponging between
// keys and keys_alt (vals/vals_alt optional).  Returns 0 if the result is in keys, 1 if in
// keys_alt.  counts holds radix_count_ints(n) ints, tile_buf scan_tile_count of that.
int64_t radix_count_ints(int64_t n);
int radix_sort_u32(uint32_t* keys, uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n, int begin_bit,
                   int end_bit, int* counts, int* tile_buf, hipStream_t s);
int radix_sort_u64(uint64_t* keys, uint64_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n, int begin_bit,
                   int end_bit, int* counts, int* tile_buf, hipStream_t s);
// Digit width of the first pass of a sort over `bits` bits (the remainder; later passes take 8).
int radix_first_digit_bits(int bits);
// radix_sort_u32 over bits [0, end_bit) of (key, input position) pairs: the values are not read
// but generated (vals_alt / vals receive the positions), and the first pass's per-tile digit
// counts are already in counts (digit-major, tiles of kRadixTile keys, written by the kernel
// that made the keys), so the sort starts with that pass's digit scan.
int radix_sort_u32_positions(uint32_t* keys, uint32_t* keys_alt, uint32_t* vals, uint32_t* vals_alt, int64_t n,
                             int end_bit, int* counts, hipStream_t s);

}  // namespace gvdbx
