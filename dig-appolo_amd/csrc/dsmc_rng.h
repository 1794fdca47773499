// Counter-based random numbers for the DSMC engine: Philox4x32-10
// (J. K. Salmon, M. A. Moraes, R. O. Dror, D. E. Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11).  The same header is compiled for the
// host (C++ restatement, host set-up code) and for gfx950, so a stream keyed
// by the same (key, counter) yields identical draws on both sides.
//
// The reference draws from one thread_local std::mt19937 seeded from
// std::random_device (src/MeshparticalInitial.h:194-203), so its draw order is
// neither reproducible nor parallel.  Here every consumer owns a stream keyed
// by what it is about: (particle id, step, wall-hit #) for wall reflections,
// (patch, step, new-particle #) for inflow emission, (cell, step) for
// collision pairs.  The draw *sequence inside* each consumer follows the
// reference's call order.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define DSMC_HD __host__ __device__ inline
#else
#define DSMC_HD inline
#endif

namespace dsmc {

enum RngPurpose : uint32_t {
  RNG_WALL = 1u,        // diffuse wall reflection during the move stage
  RNG_WALL_EMIT = 2u,   // diffuse wall reflection while tracing a new inflow particle
  RNG_EMIT = 3u,        // inflow particle sampling
  RNG_COLLIDE = 4u,     // NTC pair selection and collision draws, per cell
  RNG_INIT = 5u,        // initial particle fill, per cell
  RNG_INIT_COUNT = 6u,  // random rounding of the initial per-cell count
  RNG_GSIS = 7u,        // continuum coupling: count rounding (sub 0), cell update (sub 1)
};

struct Philox4 {
  uint32_t v[4];
};

DSMC_HD void mulhilo32(uint32_t a, uint32_t b, uint32_t& hi, uint32_t& lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// Philox4x32 with 10 rounds; multipliers and Weyl key increments are the
// published ones.
DSMC_HD Philox4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                              uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#if defined(__HIP_DEVICE_COMPILE__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo32(M0, c0, hi0, lo0);
    mulhilo32(M1, c2, hi1, lo1);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n1 = lo1;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    const uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0; k1 += W1;
  }
  Philox4 out;
  out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
  return out;
}

// A sequential stream: counter (a, b, c, block) under key (k0, k1).  Each
// Philox call yields four 32-bit words; `block` advances when they are used.
struct Stream {
  uint32_t k0, k1, a, b, c, block;
  uint32_t buf[4];
  uint32_t used;

  DSMC_HD void init(uint64_t key, uint32_t a_, uint32_t b_, uint32_t purpose, uint32_t sub) {
    k0 = (uint32_t)key;
    k1 = (uint32_t)(key >> 32);
    a = a_;
    b = b_;
    c = (purpose << 24) ^ (sub & 0x00FFFFFFu);
    block = 0u;
    used = 4u;
  }
  DSMC_HD uint32_t next_u32() {
    if (used == 4u) {
      const Philox4 r = philox4x32_10(a, b, c, block, k0, k1);
      buf[0] = r.v[0]; buf[1] = r.v[1]; buf[2] = r.v[2]; buf[3] = r.v[3];
      ++block;
      used = 0u;
    }
    // pop the front word; constant indices only, so the buffer stays in
    // registers on the GPU
    const uint32_t x = buf[0];
    buf[0] = buf[1];
    buf[1] = buf[2];
    buf[2] = buf[3];
    ++used;
    return x;
  }
  // Uniform double in (0, 1]: 53 random bits, offset by one unit so that
  // log() of a draw is always finite.  Exact integer arithmetic, so the
  // result is the same on every IEEE-754 machine.
  DSMC_HD double uniform() {
    const uint32_t x = next_u32();
    const uint32_t y = next_u32();
    const uint64_t m = ((uint64_t)(x >> 5) << 26) | (uint64_t)(y >> 6);
    return (double)(m + 1u) * 1.1102230246251565404236316680908203125e-16;  // 2^-53
  }
  // Uniform integer in [0, n), n >= 1 (Lemire's multiply-shift, no rejection;
  // bias <= n / 2^32).
  DSMC_HD uint32_t below(uint32_t n) {
    return (uint32_t)(((uint64_t)next_u32() * (uint64_t)n) >> 32);
  }
};

}  // namespace dsmc
