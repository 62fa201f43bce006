/*
 * ptk_prng.h -- the one pseudo-random generator used for every synthetic input.
 *
 * Counter based: value n of stream s under seed k is a pure function of (k, s, n), so a host
 * loop, a device kernel and a multi-threaded fill produce the same bits in any order.  The
 * mixer is SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number generators",
 * OOPSLA 2014): golden-gamma increment, then two xor-shift-multiply rounds.
 *
 * Usable from C, C++ and HIP device code.
 */
#ifndef PTK_PRNG_H
#define PTK_PRNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define PTK_PRNG_FN __host__ __device__ static inline
#else
#define PTK_PRNG_FN static inline
#endif

#ifdef __cplusplus
extern "C" {
#endif

PTK_PRNG_FN uint64_t ptk_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* 64 random bits for element n of stream `stream` under `seed`. */
PTK_PRNG_FN uint64_t ptk_rand_u64(uint64_t seed, uint64_t stream, uint64_t n) {
    return ptk_splitmix64(ptk_splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) + n);
}

/* Uniform on [-1, 1): the top 24 bits as k/2^23 - 1, exact in binary32. */
PTK_PRNG_FN float ptk_rand_sym_f32(uint64_t seed, uint64_t stream, uint64_t n) {
    uint32_t k = (uint32_t)(ptk_rand_u64(seed, stream, n) >> 40);
    return (float)k * (1.0f / 8388608.0f) - 1.0f;
}

/* Uniform on [0, 1): the top 24 bits as k/2^24, exact in binary32. */
PTK_PRNG_FN float ptk_rand_unit_f32(uint64_t seed, uint64_t stream, uint64_t n) {
    uint32_t k = (uint32_t)(ptk_rand_u64(seed, stream, n) >> 40);
    return (float)k * (1.0f / 16777216.0f);
}

/* Uniform on [0, 1) with 53 bits, exact in binary64. */
PTK_PRNG_FN double ptk_rand_unit_f64(uint64_t seed, uint64_t stream, uint64_t n) {
    return (double)(ptk_rand_u64(seed, stream, n) >> 11) * (1.0 / 9007199254740992.0);
}

#ifdef __cplusplus
}
#endif

#endif /* PTK_PRNG_H */
