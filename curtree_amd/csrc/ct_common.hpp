// ct_common.hpp -- device helpers shared by the build and query kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ct {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Box test: closed-interval overlap; touching counts as a hit and a NaN
// coordinate makes every comparison false, i.e. "intersects".  Same predicate
// as src/rtree_bfs_frontier.cuh:40-44 / src/utility.h:16-20.  Pure compares,
// no arithmetic, so it is bit-exact by construction.
__device__ __forceinline__ bool intersect(float x1, float y1, float x2, float y2, float m1,
                                          float n1, float m2, float n2) {
  return !(x1 > m2 || x2 < m1 || y1 > n2 || y2 < n1);
}

// ---------------------------------------------------------------------------
// MBR min / max.  NaN operands are ignored (as fminf/fmaxf do); -0 orders below
// +0 and two NaNs give the canonical quiet NaN, which makes both operations
// commutative and associative on bit patterns, so any reduction tree gives the
// same bits as a left-to-right loop.
__device__ __forceinline__ float canon_nan() { return __uint_as_float(0x7FC00000u); }

__device__ __forceinline__ float mbr_min(float a, float b) {
  if (a != a) return (b != b) ? canon_nan() : b;
  if (b != b) return a;
  if (a < b) return a;
  if (b < a) return b;
  return (__float_as_uint(a) >> 31) ? a : b;
}

__device__ __forceinline__ float mbr_max(float a, float b) {
  if (a != a) return (b != b) ? canon_nan() : b;
  if (b != b) return a;
  if (a > b) return a;
  if (b > a) return b;
  return (__float_as_uint(a) >> 31) ? b : a;
}

// Order-preserving map float -> uint32 (for atomic min/max of non-NaN values;
// -0 maps below +0).
// (negative: all bits flipped; otherwise the sign bit set)
__device__ __forceinline__ uint32_t f2ord(float f) {
  const uint32_t u = __float_as_uint(f);
  return u ^ ((uint32_t)((int32_t)u >> 31) | 0x80000000u);
}
__device__ __forceinline__ float ord2f(uint32_t u) {
  return __uint_as_float(u ^ ((uint32_t)((int32_t)~u >> 31) | 0x80000000u));
}

// ---------------------------------------------------------------------------
// Morton codes.  Bit spreading by binary magic numbers: x lands in the even
// bits, y in the odd bits (the z_order convention of src/z_order.cuh:67-69).
__device__ __forceinline__ uint32_t spread16(uint32_t x) {
  x &= 0xFFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}

__device__ __forceinline__ uint64_t spread32(uint32_t v) {
  uint64_t x = v;
  x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
  x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
  x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
  x = (x | (x << 2)) & 0x3333333333333333ull;
  x = (x | (x << 1)) & 0x5555555555555555ull;
  return x;
}

// (ushort)f of src/rtree_bulkload_thrust.cuh:55-56: truncation toward zero,
// saturated to [0, 65535], NaN -> 0.
__device__ __forceinline__ uint32_t f2u16_sat(float v) {
  if (!(v > 0.0f)) return 0u;
  if (v >= 65535.0f) return 65535u;
  return (uint32_t)v;
}

__device__ __forceinline__ uint32_t q32(float c, float lo, float s) {
  float t = (c - lo) * s;
  if (!(t > 0.0f)) return 0u;
  if (t >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)t;
}

// ---------------------------------------------------------------------------
// Wave primitives.
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
  return (1ull << lane_id()) - 1ull;
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }

// Device-coherent (agent scope) loads / stores for inter-workgroup
// communication across the XCDs' separate L2s.
template <class T>
__device__ __forceinline__ T ld_agent(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_agent(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Chained-scan status words for 64-bit prefix sums (query kernels).  One word
// per tile: [tag:8][value:56].  tag = 2*epoch+1 means "tile aggregate",
// tag = 2*epoch+2 means "inclusive prefix"; any other tag is "not ready".  The
// epoch changes every launch, so the status array needs zeroing only when the
// epoch counter wraps.
constexpr int kEpochs = 127;
__device__ __forceinline__ uint64_t st_pack(uint32_t tag, uint64_t v) {
  return ((uint64_t)tag << 56) | (v & 0x00FFFFFFFFFFFFFFull);
}
__device__ __forceinline__ uint32_t st_tag(uint64_t w) { return (uint32_t)(w >> 56); }
__device__ __forceinline__ uint64_t st_val(uint64_t w) { return w & 0x00FFFFFFFFFFFFFFull; }

// Wave-parallel decoupled look-back for tile `tile` (called by one full wave):
// returns the exclusive prefix of all tiles before it.  Lane i inspects tile
// (j - i); the scan stops at the nearest inclusive prefix.
__device__ __forceinline__ uint64_t lookback64(uint64_t* status, int64_t tile, uint32_t epoch) {
  const uint32_t tagA = 2u * epoch + 1u, tagP = 2u * epoch + 2u;
  uint64_t sum = 0;
  int64_t j = tile - 1;
  const int lane = lane_id();
  while (j >= 0) {
    int64_t idx = j - lane;
    uint64_t w = 0;
    uint32_t tg = tagP;
    if (idx >= 0) {
      do {
        w = ld_agent(status + idx);
        tg = st_tag(w);
      } while (tg != tagA && tg != tagP);
    } else {
      w = st_pack(tagP, 0);
    }
    uint64_t pm = ballot(tg == tagP);
    uint64_t v = st_val(w);
    if (pm) {
      int first = __ffsll((unsigned long long)pm) - 1;  // nearest inclusive prefix
      if (lane > first) v = 0;
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      sum += v;
      break;
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    sum += v;
    j -= kWave;
  }
  return sum;
}

}  // namespace ct
