// splat.hpp -- launchers of the trilinear splat and the splat / sampling adjoints (splat.hip).
#pragma once
#include "kernels.hpp"

namespace gvdbx {

// Sort key of every point, key = slot * 27 + region of the home brick (sentinel = points in no
// brick), with the first radix pass's tile counts over `bits` key bits (radix_sort_u32_positions
// then sorts them; counts holds radix_count_ints(n) ints).
void splat_keys(const DevTree& T, const float* q, int64_t n, const int* leaf_slot, int lgres, uint32_t sentinel,
                uint32_t* keys, int bits, int* counts, hipStream_t s);
// payload[j] = (q, m) of point perm[j] (m = 1 without masses), by way of inv (n words of
// scratch: the inverse permutation)
void splat_place(const float* q, const float* mass, int64_t n, const uint32_t* perm, uint32_t* inv, float4* payload,
                 hipStream_t s);
bool splat_brick_supported(int lgres);
// start: [27 * nslots + 1] key starts (exclusive scan of the counts)
void splat_brick(int64_t nslots, const int* start, const float4* payload, const int* nbr, const int4* slot_pos,
                 int lgres, int apron, float* chan, int accum, hipStream_t s);
// order: a permutation of [0, n) to visit the points in (locality only), or null
void splat_backward(const DevTree& T, const int* leaf_slot, const int* nbr, int lgres, int apron, const float* chan,
                    const float* q, const float* mass, int64_t n, const uint32_t* order, float* dq, float* dm,
                    hipStream_t s);
void sample_backward(const DevTree& T, const int* leaf_slot, int res0, int apron, const float* chan, const float* q,
                     const float* dout, int64_t n, float* dq, hipStream_t s);
// ScatterDensity: chan (padded bricks of nslots slots) += the points' (1 - c)^3 contributions,
// summed exactly as 64-bit integers of scale 2^shift in acc (nslots * br^3 entries, cleared here).
// Colour (cmode 1: average per voxel, cbuf = 4 x u64 per interior voxel; 2: the colour of the
// voxel's highest-index point, cbuf = u32 per interior voxel; cleared here) into the uchar4
// channel cchan, for the interior voxels the points reached.
void scatter_density(const DevTree& T, const PointsView& P, const int* leaf_slot, int lgres, int apron, float r,
                     int expand, int shift, int64_t nslots, unsigned long long* acc, float* chan, int cmode,
                     void* cbuf, uint32_t* cchan, hipStream_t s);

}  // namespace gvdbx
