// crt_rng.hpp -- counter-based random numbers (no generator state).
//
// u01(seed, a, b, c) hashes a 4-tuple with the splitmix64 finaliser chain and
// returns 53 random bits as a double in [0, 1).  Keys used by the tracer:
//   a = key_offset + pixel (or beam), b = sample (1-based) or sub-ray,
//   c = 64 * bounce + dim, dims: 0,1 pixel jitter; 2+2l, 3+2l area-light
//   sample of light l; 62, 63 cosine (BRDF) sample; 0,1 LiDAR sub-ray offsets.
// Any result therefore depends only on its key, never on launch order or
// batching.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace crt {

constexpr unsigned RNG_JITTER = 0u;
constexpr unsigned RNG_LIGHT = 2u;
constexpr unsigned RNG_BRDF = 62u;

__host__ __device__ inline uint64_t rng_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The first three rounds depend on (seed, a, b) only: a caller drawing
// several numbers for one pixel and sample computes that key once and pays
// one round per number (a 64-bit multiply is several VALU instructions).
// rng_u01(seed, a, b, c) == rng_u01k(rng_key(seed, a, b), c), bit for bit.
__host__ __device__ inline uint64_t rng_key(uint64_t seed, uint64_t a, uint64_t b) {
    uint64_t h = rng_mix64(seed + 0x9E3779B97F4A7C15ull);
    h = rng_mix64(h ^ a);
    return rng_mix64(h ^ (b + 0x632BE59BD9B4E019ull));
}

__host__ __device__ inline double rng_u01k(uint64_t key, uint64_t c) {
    const uint64_t h = rng_mix64(key ^ (c * 0x8CB92BA72F3D8DD7ull + 1ull));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline double rng_u01(uint64_t seed, uint64_t a, uint64_t b, uint64_t c) {
    return rng_u01k(rng_key(seed, a, b), c);
}

__host__ __device__ inline uint64_t rng_dim(int bounce, unsigned dim) { return (uint64_t)bounce * 64u + dim; }

}  // namespace crt
