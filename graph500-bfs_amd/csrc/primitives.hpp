// primitives.hpp -- device-wide scan and LSD radix sort (gfx950, wave64).
//
// Used by the CSR build (vertex relabel by degree, edge sort + dedupe) and by
// the top-down frontier expansion (prefix sum of frontier degrees).
#pragma once

#include "common.hpp"

namespace g500 {

// Scratch for exclusive scans of up to `cap` elements (all recursion levels).
struct ScanWorkspace {
  DevBuf<uint64_t> buf;
  int64_t cap = 0;
  void reserve(int64_t n);
};

// out[i] = sum(in[0..i)) for i in [0, n]; out must hold n+1 entries
// (out[n] = total).  in and out may alias.  InT is uint32_t or uint64_t.
template <typename InT>
void exclusive_scan(const InT* in, uint64_t* out, int64_t n, ScanWorkspace& ws,
                    hipStream_t s);

// Scratch for radix sorts of up to `cap` keys.
struct RadixWorkspace {
  DevBuf<uint32_t> counts;  // [256][tiles]
  DevBuf<uint64_t> offsets; // [256][tiles] + 1
  ScanWorkspace scan;
  int64_t cap = 0;
  void reserve(int64_t n);
};

// Sorts keys[0..n) ascending on bits [0, bits).  `alt` must hold n keys.
// Returns the buffer that holds the result (keys or alt).
uint64_t* radix_sort_u64(uint64_t* keys, uint64_t* alt, int64_t n, int bits,
                         RadixWorkspace& ws, hipStream_t s);

constexpr int kRadixItems = 32;              // keys per thread per tile
constexpr int kRadixThreads = 256;           // 4 waves
constexpr int kRadixTile = kRadixItems * kRadixThreads;

}  // namespace g500
