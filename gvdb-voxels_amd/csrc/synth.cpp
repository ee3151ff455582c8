// synth.cpp -- seeded synthetic point sets, the same bits in C++ and Python (gvdbx/synth.py).
//
// Draw k of stream `seed` is the SplitMix64 finalizer of seed * 0xD1B54A32D192ED03 +
// (k + 1) * 0x9E3779B97F4A7C15 (wrapping 64-bit arithmetic); a uniform is its top 53 bits times
// 2^-53, a normal the sum of 12 uniforms minus 6 (Irwin-Hall), added in order.  Only exactly
// rounded IEEE operations follow (+, -, *, /, sqrt, the FP64 -> FP32 cast), so a port that does
// the same operations in the same order produces the same floats.  Built without contraction
// (-ffp-contract=off).
#include <cmath>
#include <cstdint>

#include "../../include/gvdbx.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

inline uint64_t draw(uint64_t seed, uint64_t k) {
  return mix64(seed * 0xD1B54A32D192ED03ull + (k + 1) * 0x9E3779B97F4A7C15ull);
}

inline double uniform(uint64_t seed, uint64_t k) { return (double)(draw(seed, k) >> 11) * 0x1.0p-53; }

// normal from draws k .. k + 11
inline double normal(uint64_t seed, uint64_t k) {
  double s = 0.0;
  for (int j = 0; j < 12; ++j) s += uniform(seed, k + j);
  return s - 6.0;
}

}  // namespace

extern "C" {

int gvdbx_synth_uniform(uint64_t seed, int64_t n, const float lo[3], const float hi[3], float* out) {
  if (n < 0 || !lo || !hi || (n > 0 && !out)) return GVDBX_ERR_ARG;
  for (int64_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      const double u = uniform(seed, (uint64_t)(3 * i + a));
      out[3 * i + a] = (float)((double)lo[a] + ((double)hi[a] - (double)lo[a]) * u);
    }
  return GVDBX_OK;
}

int gvdbx_synth_sphere(uint64_t seed, int64_t n, const float center[3], float radius, float sigma, float* out) {
  if (n < 0 || !center || (n > 0 && !out)) return GVDBX_ERR_ARG;
  for (int64_t i = 0; i < n; ++i) {
    const uint64_t k = (uint64_t)i * 48;
    const double d0 = normal(seed, k), d1 = normal(seed, k + 12), d2 = normal(seed, k + 24);
    const double len = std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    const double rho = (double)radius + (double)sigma * normal(seed, k + 36);
    const double s = len > 0.0 ? rho / len : 0.0;
    out[3 * i + 0] = (float)((double)center[0] + d0 * s);
    out[3 * i + 1] = (float)((double)center[1] + d1 * s);
    out[3 * i + 2] = (float)((double)center[2] + d2 * s);
  }
  return GVDBX_OK;
}

int gvdbx_synth_lattice(uint64_t seed, const int dims[3], float spacing, float jitter, float vel_sigma, float* pos,
                        float* vel) {
  if (!dims || dims[0] < 0 || dims[1] < 0 || dims[2] < 0) return GVDBX_ERR_ARG;
  const int64_t n = (int64_t)dims[0] * dims[1] * dims[2];
  if (n > 0 && !pos) return GVDBX_ERR_ARG;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t ix = i / ((int64_t)dims[1] * dims[2]), iy = (i / dims[2]) % dims[1], iz = i % dims[2];
    const int64_t c[3] = {ix, iy, iz};
    const uint64_t k = (uint64_t)i * 39;
    for (int a = 0; a < 3; ++a) {
      const double u = uniform(seed, k + a);
      pos[3 * i + a] = (float)(((double)c[a] + 0.5) * (double)spacing + (u - 0.5) * (double)jitter);
      if (vel) vel[3 * i + a] = (float)((double)vel_sigma * normal(seed, k + 3 + 12 * a));
    }
  }
  return GVDBX_OK;
}

}  // extern "C"
