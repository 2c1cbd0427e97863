/*
 * bmx_rng.h -- counter-based uniform random numbers, the same on the host
 * and on gfx950.
 *
 * The reference draws from AMReX's per-thread engines inside
 * ParallelForRNG (src/des/bmx_split_particles.cpp:43-58, 83-160), so its
 * sequence depends on the launch and cannot be reproduced.  Here every draw
 * is a pure function of (seed, call, particle id, particle cpu, draw index):
 * the same particle gets the same numbers on the GPU and in the oracle, in
 * any order.  Mixing: the splitmix64 output function (Steele, Lea and
 * Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014)
 * applied over the counters, twice; 53 bits to a double in [0, 1).
 */
#ifndef BMX_RNG_H
#define BMX_RNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define BMX_RNG_FN __host__ __device__ static inline
#else
#define BMX_RNG_FN static inline
#endif

BMX_RNG_FN uint64_t bmx_mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

/* draw-th uniform in [0, 1) of particle (id, cpu) in call 'call' */
BMX_RNG_FN double bmx_uniform(uint64_t seed, uint32_t call, int32_t id,
                              int32_t cpu, uint32_t draw) {
  uint64_t h = bmx_mix64(seed ^ ((uint64_t)call << 32 | (uint64_t)draw));
  h = bmx_mix64(h ^ ((uint64_t)(uint32_t)id << 32 | (uint64_t)(uint32_t)cpu));
  return (double)(h >> 11) * 0x1p-53;
}

#endif
