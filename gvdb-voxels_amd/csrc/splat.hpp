// splat.hpp -- launchers of the trilinear splat and the splat / sampling adjoints (splat.hip).
#pragma once
#include "kernels.hpp"

namespace gvdbx {

// key = slot * 27 + region of the home brick, sentinel for points in no brick
void splat_keys(const DevTree& T, const float* q, int64_t n, const int* leaf_slot, int lgres, uint32_t sentinel,
                uint32_t* keys, uint32_t* vals, hipStream_t s);
bool splat_brick_supported(int lgres);
// start: [27 * nslots + 2] range starts of the sorted keys; mass may be null (unit masses)
void splat_brick(int64_t nslots, const int* start, const uint32_t* svals, const float* q, const float* mass,
                 const int* nbr, const int4* slot_pos, int lgres, int apron, float* chan, int accum, hipStream_t s);
void splat_backward(const DevTree& T, const int* leaf_slot, const int* nbr, int lgres, int apron, const float* chan,
                    const float* q, const float* mass, int64_t n, float* dq, float* dm, hipStream_t s);
void sample_backward(const DevTree& T, const int* leaf_slot, int res0, int apron, const float* chan, const float* q,
                     const float* dout, int64_t n, float* dq, hipStream_t s);

}  // namespace gvdbx
