// device_utils.h -- CDNA device helpers: 64-lane wavefront reductions and the
// XCD-aware block remapping used by the streaming kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace afem {

constexpr int kWave = 64;

template <class T>
__device__ __forceinline__ T wave_reduce_sum(T v)
{
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1)
    v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum over groups of G consecutive lanes (G a power of two <= 64); every lane
// of the group receives the group sum.
template <int G, class T>
__device__ __forceinline__ T group_reduce_sum(T v)
{
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
    v += __shfl_xor(v, off, kWave);
  return v;
}

// MI355X dispatches consecutive workgroups round-robin over its 8 XCDs, each
// with a private L2.  This maps the hardware block id to a logical tile id so
// that every XCD works on one contiguous range of tiles (neighbouring rows
// share coordinates and vector entries through that XCD's L2).  Bijective on
// [0, nblocks).
constexpr int kNumXcd = 8;

__device__ __forceinline__ uint32_t xcd_tile(uint32_t bid, uint32_t nblocks)
{
  uint32_t xcd = bid % kNumXcd;
  uint32_t local = bid / kNumXcd;
  uint32_t q = nblocks / kNumXcd;
  uint32_t rem = nblocks % kNumXcd;
  uint32_t start = xcd * q + (xcd < rem ? xcd : rem);
  return start + local;
}

} // namespace afem
