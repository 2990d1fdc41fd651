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

// Decoupled look-back (Merrill & Garland 2016) for one value. `status[t]`
// already holds this tile's aggregate; walks predecessors until an inclusive
// prefix is found and returns the exclusive prefix of tile t.
__device__ __forceinline__ uint32_t lookback_exclusive(uint32_t* status, uint32_t t, uint32_t* err) {
  uint32_t excl = 0;
  int64_t j = int64_t(t) - 1;
  uint32_t spins = 0;
  while (j >= 0) {
    const uint32_t s = status_load(status + j);
    const uint32_t flag = s & kFlagMask;
    if (flag == 0u) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > kSpinLimit) {
        atomicOr(err, kErrLookbackSpin);
        break;
      }
      continue;
    }
    excl += s & kValueMask;
    if (flag == kFlagInclusive) break;
    --j;
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
