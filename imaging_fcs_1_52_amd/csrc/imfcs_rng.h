// Counter-based random numbers for device code: every draw is a hash of
// (seed, stream index, step, purpose, k), so results do not depend on the
// launch shape or the order in which threads run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace imfcs {

// SplitMix64 finalizer (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014) as the mixing function.
__device__ __forceinline__ uint64_t fmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// key of the random numbers of (index e, step t, purpose)
__device__ __forceinline__ uint64_t rkey(uint64_t seed, uint64_t e, uint64_t t, uint64_t purpose) {
    const uint64_t h = fmix64(seed ^ (e * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull));
    return fmix64(h ^ (t * 0xD1B54A32D192ED03ull + purpose * 0x8CB92BA72F3D8DD7ull));
}

// k-th uniform of a key in (0, 1]
__device__ __forceinline__ float u01(uint64_t key, uint32_t k) {
    const uint64_t z = fmix64(key + (uint64_t)(k + 1) * 0x9E3779B97F4A7C15ull);
    return ((float)(z >> 40) + 1.0f) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ double u01d(uint64_t key, uint32_t k) {
    const uint64_t z = fmix64(key + (uint64_t)(k + 1) * 0x9E3779B97F4A7C15ull);
    return ((double)(z >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

// standard normal from uniforms k, k+1 (Box-Muller)
__device__ __forceinline__ float nrm(uint64_t key, uint32_t k) {
    const float r = sqrtf(-2.f * logf(u01(key, k)));
    return r * cosf(6.2831853071795865f * u01(key, k + 1));
}

}  // namespace imfcs
