// mrg.hpp -- L'Ecuyer MRG (k = 5, m = 2^31-1) used by the Graph500 generator,
// as host + device code.
//
// One step maps the state (z1..z5) to (x*z1 + y*z5, z1, z2, z3, z4) mod m with
// x = 107374182, y = 104480 (generator/splittable_mrg.c:17-31, 210-216).  A
// skip by e steps multiplies the state by A^e; this file keeps A^(b * 256^k)
// for every byte position k and byte value b as dense 5x5 matrices computed
// at start-up, so a skip is one matrix-vector product per non-zero byte of
// the exponent (the reference's byte-indexed idea, splittable_mrg.c:224-241,
// with our own matrix layout and table construction).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace g500 {

constexpr uint32_t kMrgP = 0x7FFFFFFFu;
constexpr uint32_t kMrgX = 107374182u;
constexpr uint32_t kMrgY = 104480u;

struct MrgState {
  uint32_t z[5];
};

struct Mat5 {
  uint32_t a[25];  // row-major
};

// x < 2^62 -> x mod (2^31 - 1), using 2^31 == 1 (mod m).
__host__ __device__ inline uint32_t mrg_reduce(uint64_t x) {
  uint64_t r = (x & kMrgP) + (x >> 31);
  r = (r & kMrgP) + (r >> 31);
  return (uint32_t)(r >= kMrgP ? r - kMrgP : r);
}

__host__ __device__ inline uint32_t mrg_next(MrgState& s) {
  // x*z1 < 2^58 and y*z5 < 2^48, so the sum stays below 2^62.
  const uint32_t nz =
      mrg_reduce((uint64_t)kMrgX * s.z[0] + (uint64_t)kMrgY * s.z[4]);
  s.z[4] = s.z[3];
  s.z[3] = s.z[2];
  s.z[2] = s.z[1];
  s.z[1] = s.z[0];
  s.z[0] = nz;
  return nz;
}

__host__ __device__ inline void mat5_apply(const Mat5& m, MrgState& s) {
  uint32_t o[5];
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    uint64_t acc = 0;
#pragma unroll
    for (int k = 0; k < 5; ++k)
      acc += mrg_reduce((uint64_t)m.a[i * 5 + k] * s.z[k]);
    o[i] = mrg_reduce(acc);
  }
#pragma unroll
  for (int i = 0; i < 5; ++i) s.z[i] = o[i];
}

// generator/graph_generator.hpp:492-498: two user seeds -> five state words.
__host__ __device__ inline MrgState mrg_seed_from_user(uint64_t s1,
                                                       uint64_t s2) {
  MrgState st;
  st.z[0] = (uint32_t)((s1 & 0x3FFFFFFF) + 1);
  st.z[1] = (uint32_t)(((s1 >> 30) & 0x3FFFFFFF) + 1);
  st.z[2] = (uint32_t)((s2 & 0x3FFFFFFF) + 1);
  st.z[3] = (uint32_t)(((s2 >> 30) & 0x3FFFFFFF) + 1);
  st.z[4] = (uint32_t)(((s2 >> 60) << 4) + (s1 >> 60) + 1);
  return st;
}

// Host-side table: byte position k (0..23) x byte value b (0..255).
struct MrgTable {
  Mat5 t[24][256];
};

// Builds the table once (thread-safe, lazily) and returns it.
const MrgTable& mrg_table();

// state <- A^(hi*2^128 + mid*2^64 + lo) * state  (host).
void mrg_skip_host(MrgState& s, uint64_t hi, uint64_t mid, uint64_t lo);

// splittable_mrg.c:337-343: two draws combined into a double in [0,1).
double mrg_get_double_host(MrgState& s);

// graph_generator.hpp:503-520.
void make_random_numbers_host(int64_t n, uint64_t s1, uint64_t s2,
                              int64_t position, double* out);

// graph_generator.hpp:528-538: the two scramble words of a seed pair.
void scramble_values_host(uint64_t s1, uint64_t s2, uint64_t out[2]);

}  // namespace g500
