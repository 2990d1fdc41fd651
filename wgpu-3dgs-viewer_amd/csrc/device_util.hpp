// Wavefront (64-lane) and cross-workgroup primitives for gfx950.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "types.hpp"

namespace gs3d {

constexpr int kWave = 64;

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ uint32_t mask_rank(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Device-scope (all XCDs) relaxed load/store of a self-contained status word.
__device__ __forceinline__ uint32_t status_load(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void status_store(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Spin limit for look-back waits: a predecessor that holds a ticket is
// resident and publishes within microseconds; the bound only turns a logic
// error into a reported error instead of a hang.
constexpr uint32_t kSpinLimit = 1u << 20;

// Sum over the wave's 64 lanes (every lane active), in every lane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Decoupled look-back (Merrill & Garland 2016) for one value, run by a whole
// wave (all 64 lanes active; the result in every lane). `status[t]` already
// holds this tile's aggregate; returns the exclusive prefix of tile t. Each
// step loads the 64 nearest unexamined predecessors' words at once (one
// round trip instead of one per predecessor), adds the aggregates up to the
// nearest inclusive prefix and stops there or, when a predecessor has not
// published yet, adds those before it and continues from it.
__device__ __forceinline__ uint32_t lookback_exclusive_wave(uint32_t* status, uint32_t t, uint32_t* err) {
  const uint32_t lane = lane_id();
  uint32_t excl = 0, spins = 0;
  int64_t hi = int64_t(t) - 1;
  while (hi >= 0) {
    const int64_t j = hi - int64_t(lane);
    // below tile 0: an inclusive prefix of 0
    const uint32_t s = j >= 0 ? status_load(status + j) : kFlagInclusive;
    const uint32_t flag = s & kFlagMask;
    const uint64_t pending = __ballot(flag == 0u);
    const uint64_t incl = __ballot(flag == kFlagInclusive);
    const uint32_t first_pending = pending ? uint32_t(__builtin_ctzll(pending)) : 64u;
    const uint32_t first_incl = incl ? uint32_t(__builtin_ctzll(incl)) : 64u;
    if (first_incl < first_pending) {
      excl += wave_sum(lane <= first_incl ? (s & kValueMask) : 0u);
      break;
    }
    excl += wave_sum(lane < first_pending ? (s & kValueMask) : 0u);
    hi -= int64_t(first_pending);
    if (first_pending == 0u) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kSpinLimit) {
        if (lane == 0) atomicOr(err, kErrLookbackSpin);
        break;
      }
    }
  }
  return excl;
}

// Ticket: the order in which workgroups start, so that every predecessor a
// workgroup waits on is already resident.
__device__ __forceinline__ uint32_t take_ticket(uint32_t* ticket, uint32_t* s_slot) {
  __syncthreads();  // every lane has read the previous ticket from s_slot
  if (threadIdx.x == 0) *s_slot = atomicAdd(ticket, 1u);
  __syncthreads();
  const uint32_t t = *s_slot;
  return t;
}

}  // namespace gs3d
