// rng.h -- counter-based random numbers and the source distributions shared
// by the device particle sources (single domain and ensemble).
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): a stream is keyed by a 64-bit key and counted by
// (c0, c1, tag, block), four 32-bit words per block, two words per uniform
// double (53 random bits).  Box-Muller as Tools.f90:30-48, the drifting
// Maxwellian of Tools.f90:1-28 and the dt/2 rotation with norm restore of
// ParticlesInitialize.f90:76-84.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace picard {

struct Philox {
  uint32_t c0, c1, c2, c3;
};

__device__ __forceinline__ void philox_round(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3,
                                             uint32_t k0, uint32_t k1) {
  const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
  const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
  const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
  const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
  const uint32_t n0 = hi1 ^ c1 ^ k0;
  const uint32_t n2 = hi0 ^ c3 ^ k1;
  c0 = n0;
  c1 = lo1;
  c2 = n2;
  c3 = lo0;
}

// Philox4x32-10 of counter (c0..c3) with a 64-bit key.
__device__ __forceinline__ Philox philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                         uint64_t key) {
  uint32_t k0 = (uint32_t)key, k1 = (uint32_t)(key >> 32);
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    philox_round(c0, c1, c2, c3, k0, k1);
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return {c0, c1, c2, c3};
}

// uniform double in [0, 1) from 53 random bits
__device__ __forceinline__ double u01(uint32_t a, uint32_t b) {
  const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);
  return (double)m * (1.0 / 9007199254740992.0);
}

// A per-particle generator: successive 4-word blocks of one counter stream.
struct Stream {
  uint64_t key;
  uint32_t slot_lo, slot_hi, tag;
  uint32_t block = 0;
  uint32_t w[4];
  int used = 4;
  __device__ uint32_t next() {
    if (used == 4) {
      const Philox r = philox(slot_lo, slot_hi, tag, block++, key);
      w[0] = r.c0; w[1] = r.c1; w[2] = r.c2; w[3] = r.c3;
      used = 0;
    }
    return w[used++];
  }
  __device__ double uniform() {
    const uint32_t a = next();
    const uint32_t b = next();
    return u01(a, b);
  }
  // Box-Muller as Tools.f90:30-48, r2 > 0
  __device__ double normal() {
    const double r1 = uniform();
    double r2 = uniform();
    while (!(r2 > 0.0)) r2 = uniform();
    return sqrt(-2.0 * log(r2)) * sin(6.2831853071795864769252867665590 * r1);
  }
};

struct PlasmaConst {
  int32_t species;
  int32_t ipp[8], epp[8], hpp[8];
  int32_t group[8];            // 1 + epp + hpp
  int32_t per_cell;            // particles per cell over species first..S
  int32_t first_species;
  double mass[8], kelvin[8], v0[8][3];
  double ve[3], tB[8][3];
};

__device__ inline void maxwellian(Stream &st, const PlasmaConst &pc, int s_vel, int s_bulk, double v[3]) {
  const double T = pc.kelvin[s_vel], m = pc.mass[s_vel];
  if (T > 0.0 && m > 0.0) {
    const double a = m / (2.0 * 1.380650e-23 * T);
    const double pw = sqrt(1.0 / a);
    for (int c = 0; c < 3; ++c) v[c] = st.normal();
    for (int c = 0; c < 3; ++c) v[c] = v[c] * pw / sqrt(2.0) + pc.v0[s_bulk][c];
  } else {
    for (int c = 0; c < 3; ++c) v[c] = pc.v0[s_bulk][c];
  }
}

__device__ inline void half_rotate(const PlasmaConst &pc, int s, double v[3]) {
  const double *t = pc.tB[s];
  const double m1 = v[0] - pc.ve[0], m2 = v[1] - pc.ve[1], m3 = v[2] - pc.ve[2];
  double d1 = m1 + m2 * t[2] - m3 * t[1];
  double d2 = m2 + m3 * t[0] - m1 * t[2];
  double d3 = m3 + m1 * t[1] - m2 * t[0];
  const double sd = d1 * d1 + d2 * d2 + d3 * d3;
  if (sd > 0.0) {
    const double f = sqrt((m1 * m1 + m2 * m2 + m3 * m3) / sd);
    d1 *= f; d2 *= f; d3 *= f;
  }
  v[0] = d1 + pc.ve[0];
  v[1] = d2 + pc.ve[1];
  v[2] = d3 + pc.ve[2];
}

}  // namespace picard
