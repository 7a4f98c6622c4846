// Lane refill for distance walks on gfx950 (64-lane waves).
//
// Shapes whose DistanceToIn or DistanceToOut is a walk of data-dependent
// length (the Polyhedron's z segments) expose it as an ops type with
// Begin / Step / End over a per-lane State (S::DistInOps, S::DistOutOps).  In the
// one-query-per-lane form a wave runs as long as its longest walk while the
// lanes whose walks ended idle.  Here a wave owns a range of queries: a lane
// without a query takes the next unassigned one (ballot + mbcnt, no atomics)
// and runs its Begin, and lanes whose query resolves there write the result
// and take another; then every lane holding a query runs one Step.  Each
// Step thus runs on at least half of the lanes until the range runs out.  Every query
// of the range is assigned to exactly one lane once, and every walk ends, so
// the loop ends for every wave.  Results are those of S::DistanceToIn: the
// same functions run on the same values, only the order across lanes
// differs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "vg_math.hpp"

namespace vga {

// A refill pass runs Begin on the idle lanes only, so it is held back until
// at least kRefillMin lanes are idle (or all are): a pass for a few lanes
// would cost as much as one for the whole wave.
constexpr int kRefillMin = 32;

// Must be called by all 64 lanes of a wave with the same [begin, end).
// load(i, p, d, sm, r) reads query i (shape frame), or returns true with its
// result in r when it resolves it without the walk; store(i, r) writes a
// result.
template <class Ops, class PT, class Load, class Store>
__device__ __forceinline__ void WalkRefill(const PT &P, uint64_t begin, uint64_t end, Load load, Store store) {
  typename Ops::State w;
  uint64_t next = begin, qi = 0;
  bool active = false;
  for (;;) {
    for (;;) {
      const uint64_t need = __ballot(!active);
      if (need == 0 || next >= end) break;
      if (need != ~0ull && __popcll(need) < kRefillMin) break;
      if (!active) {
        const uint64_t i =
            next + __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
        if (i < end) {
          V3 p, d;
          double sm, r;
          qi = i;
          if (load(i, p, d, sm, r))
            store(i, r);
          else if (Ops::Begin(P, p, d, sm, w))
            store(i, Ops::End(P, w));
          else
            active = true;
        }
      }
      next += (uint64_t)__popcll(need);
    }
    if (__ballot(active) == 0) break;
    if (active && Ops::Step(P, w)) {
      store(qi, Ops::End(P, w));
      active = false;
    }
  }
}

}  // namespace vga
