// Counter-based random numbers for the DPD random force.
//
// The reference draws one Gaussian per interacting pair from a single
// sequential ziggurat stream (dpd_forces.f90:19, ziggurat.f90:110-147), so the
// noise depends on the pair-loop order.  Here the noise of pair (i,j) at a
// given step is a pure function of (seed, step, min(i,j), max(i,j)): any
// thread visiting the pair from either side draws the same number, which keeps
// the random pair force exactly antisymmetric.
//
// Generator: Threefry-2x32 with 13 rounds (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC11), implemented from the
// published description: Threefish-style add/rotate/xor rounds with a key
// injection every four rounds.  It uses no integer multiplies (quarter rate
// on CDNA), only full-rate adds, rotates and xors.  13 rounds is the count
// the paper finds free of BigCrush failures for this variant (its default,
// 20, adds a safety margin); the deviate is computed once per pair direction
// in the force pass, so the seven rounds saved are about 21 of the ~150 VALU
// instructions of a list entry.  The known-answer vectors of both round
// counts pin the implementation.
//
// Normal transform: Box-Muller on the two 32-bit outputs, using the top 24
// bits of each word so the uniforms are exact in float and double:
//   u1 = (w0 >> 8) * 2^-24 + 2^-25   in (0,1)
//   u2 = (w1 >> 8) * 2^-24           in [0,1)
//   xi = sqrt(-2 ln u1) * cos(2 pi u2)
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define MFA_HD __host__ __device__ __forceinline__
#else
#define MFA_HD static inline
#endif

#define MFA_TF_ROUNDS 13

MFA_HD uint32_t mfa_rotl32(uint32_t x, unsigned r) { return (x << r) | (x >> (32u - r)); }

// Threefry-2x32-R: counter (c0,c1), key (k0,k1) -> (o0,o1).
MFA_HD void mfa_threefry2x32(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1, int rounds,
                             uint32_t* o0, uint32_t* o1) {
  const unsigned rot[8] = {13u, 15u, 26u, 6u, 17u, 29u, 16u, 24u};
  uint32_t ks[3];
  ks[0] = k0;
  ks[1] = k1;
  ks[2] = 0x1BD11BDAu ^ k0 ^ k1;
  uint32_t x0 = c0 + ks[0];
  uint32_t x1 = c1 + ks[1];
  for (int r = 0; r < rounds; ++r) {
    x0 += x1;
    x1 = mfa_rotl32(x1, rot[r & 7]);
    x1 ^= x0;
    if ((r & 3) == 3) {
      const uint32_t s = (uint32_t)((r + 1) >> 2);
      x0 += ks[s % 3u];
      x1 += ks[(s + 1u) % 3u] + s;
    }
  }
  *o0 = x0;
  *o1 = x1;
}

// The two words of pair (i,j) (0-based particle indices) at `step`.
MFA_HD void mfa_pair_words(uint32_t seed, int64_t step, int32_t i, int32_t j, uint32_t* w0,
                           uint32_t* w1) {
  const uint32_t lo = (uint32_t)(i < j ? i : j);
  const uint32_t hi = (uint32_t)(i < j ? j : i);
  mfa_threefry2x32(lo, hi, seed, (uint32_t)step, MFA_TF_ROUNDS, w0, w1);
}

// The two words of particle i's deviate d (0, 1, 2: the components of the
// Langevin random force, lgv_forces.f90:25-31) at `step`.  The second counter
// word has its top bit set, which no pair counter has (particle indices stay
// below 2^27), so particle and pair streams never meet.
MFA_HD void mfa_particle_words(uint32_t seed, int64_t step, int32_t i, int d, uint32_t* w0,
                               uint32_t* w1) {
  mfa_threefry2x32((uint32_t)i, 0x80000000u | (uint32_t)d, seed, (uint32_t)step, MFA_TF_ROUNDS, w0,
                   w1);
}
