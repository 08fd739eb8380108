// Shared device and host helpers of the HIP sssp path (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "row_search.hpp"

namespace dyng_hip {

using u64 = unsigned long long;
constexpr u64 packed_inf = ~0ULL;
/// DISTANCE_INF of the reference (INT64_MAX / 4): distances at or above half of it are
/// unreachable (cpp/include/dyng/core/types.hpp:91).
constexpr int64_t sssp_infinity = INT64_MAX / 4;
/// The largest finite distance an output may hold (kernels.cuh:40).
constexpr u64 output_max_distance = static_cast<u64>(sssp_infinity / 2 - 1);

#define DYNG_DEV __device__ __forceinline__

/// Packed (distance << bits | parent) words; bits == 0 means distance-only words
/// (kernels.cuh:52-78). The all-ones parent field is "no parent"; the all-ones word is INF.
struct packing {
  int bits;
  u64 no_parent;
  DYNG_DEV u64 pack(u64 distance, int32_t parent) const {
    if (bits == 0) return distance;
    return (distance << bits) | (parent < 0 ? no_parent : static_cast<u64>(static_cast<uint32_t>(parent)));
  }
  DYNG_DEV u64 distance(u64 word) const { return word >> bits; }
  DYNG_DEV int32_t parent(u64 word) const {
    const u64 p = word & no_parent;
    return p == no_parent ? -1 : static_cast<int32_t>(p);
  }
};

// ---- agent-scope loads and stores of data other workgroups write during the same kernel ----

template <typename T>
DYNG_DEV T ld(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
DYNG_DEV void st(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

DYNG_DEV int lane() { return static_cast<int>(__lane_id()); }

/// A value that is the same on every lane (a grid-wide count read after a barrier), moved to a
/// scalar register.
DYNG_DEV int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
DYNG_DEV unsigned long long uniform(unsigned long long v) {
  const unsigned int lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(v));
  const unsigned int hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned int>(v >> 32));
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

DYNG_DEV u64 lanes_below() {
  const int l = lane();
  return l == 0 ? 0ULL : (~0ULL >> (64 - l));
}

/// Reserve one slot of a shared list per lane with `pred`: one atomicAdd per wave (the active
/// lanes with pred vote, the lowest of them reserves for all). Every active lane must call it.
DYNG_DEV int wave_append(int* counter, bool pred) {
  const u64 mask = __ballot(pred);
  if (!pred) return -1;
  const int leader = __ffsll(static_cast<long long>(mask)) - 1;
  const int rank = __popcll(mask & lanes_below());
  int base = 0;
  if (lane() == leader) base = atomicAdd(counter, __popcll(mask));
  base = __shfl(base, leader);
  return base + rank;
}

DYNG_DEV u64 wave_min(u64 v) {
  for (int off = 32; off > 0; off >>= 1) {
    const u64 o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

DYNG_DEV u64 wave_sum(u64 v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

/// Reduce v over the workgroup with op; every thread gets the result. Every thread of the
/// workgroup calls it. `part` is an LDS array of at least blockDim.x / 64 entries (reused across
/// calls: the leading barrier keeps a call from overwriting what the previous one still reads).
/// Counters that every workgroup bumps are single addresses, and device-scope atomics on one
/// address serialize (1-2 ns each), so kernels reduce per workgroup before their atomics.
template <class T, class Op>
DYNG_DEV T block_reduce(T v, T* part, Op op) {
  for (int off = 32; off > 0; off >>= 1) v = op(v, static_cast<T>(__shfl_xor(v, off)));
  const int waves = static_cast<int>((blockDim.x + 63) / 64);
  __syncthreads();
  if (lane() == 0) part[threadIdx.x / 64] = v;
  __syncthreads();
  T t = part[0];
  for (int w = 1; w < waves; ++w) t = op(t, part[w]);
  return t;
}

struct op_sum {
  template <class T>
  DYNG_DEV T operator()(T a, T b) const { return a + b; }
};
struct op_min {
  template <class T>
  DYNG_DEV T operator()(T a, T b) const { return b < a ? b : a; }
};
struct op_max {
  template <class T>
  DYNG_DEV T operator()(T a, T b) const { return b > a ? b : a; }
};


/// Grid-wide barrier of a persistent kernel whose workgroups are all resident (the grid is sized
/// to at most half of what the occupancy calculator says fits at once).
///
/// Arrivals go up a two-level tree of monotonic counters: workgroup b arrives on the counter of
/// group b % G (G = min(16, workgroups)); the last member of a group to arrive (the one whose
/// fetch-add returns a count one short of a whole round) arrives on the counter of groups, and
/// the last group releases the round by storing its number. Device-scope atomics on one address
/// serialize, so 16-way groups cut the arrival chain from one atomic per workgroup to about 16 +
/// 16. The counters never reset (the round is the count divided by the members), so no arrival
/// reads the state first and no reset store can race the next round's arrivals; the state is
/// cleared once per launch. Every counter lives on its own 128-byte line, apart from the word
/// the waiters poll. Measured on MI355X with 256 workgroups of 1024 threads: 1.9 us per barrier
/// (7.7 us with one flat counter that is read, reset and bumped).
///
/// Two flavours. `Fenced` adds an agent-scope release fence before the arrival and an acquire
/// fence after the wait (on MI355X they write back and invalidate the XCD's L2), so plain loads
/// and stores on either side are ordered. Without it the barrier only waits for the thread's own
/// memory operations to complete (s_waitcnt) before the arrival: the phases it separates must
/// access every location that another workgroup writes during the kernel with agent-scope
/// operations only (ld / st / atomics), which are coherent across the XCDs' L2s by themselves.
/// That saves the two cache walks, the bulk of a fenced barrier's cost.
///
/// Every wait is bounded: a workgroup that waits longer than `limit` ticks of the 100 MHz
/// real-time counter raises `abort`, and every workgroup that sees `abort` while waiting leaves
/// too (both are checked on the first poll and every 32nd after). The function then returns
/// false and the caller leaves the kernel, so a grid whose workgroups could not all be resident
/// ends with an error instead of spinning forever.
constexpr unsigned int kBarrierGroups = 16;

struct alignas(128) barrier_line {
  unsigned long long v;
  unsigned long long pad[15];
};

struct grid_sync_state {
  barrier_line groups;                 ///< groups arrived, over all rounds
  barrier_line round;                  ///< the last released round
  barrier_line abort;                  ///< nonzero: a wait timed out
  barrier_line group[kBarrierGroups];  ///< arrivals of each group, over all rounds
};

template <bool Fenced>
DYNG_DEV bool grid_barrier(grid_sync_state* s, unsigned long long limit) {
  __shared__ int go;
  // every wave's loads and stores have completed (at the scope they were issued with) before
  // thread 0 arrives
  if (!Fenced) __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (Fenced) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned int blocks = gridDim.x;
    const unsigned int G = blocks < kBarrierGroups ? blocks : kBarrierGroups;
    const unsigned int grp = blockIdx.x % G;
    const unsigned long long members = blocks / G + (blocks % G > grp ? 1u : 0u);
    const unsigned long long a = __hip_atomic_fetch_add(&s->group[grp].v, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long target = a / members + 1;  // the round this arrival completes
    bool last = false;
    if (a % members == members - 1) {
      const unsigned long long g = __hip_atomic_fetch_add(&s->groups.v, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      last = g % G == G - 1;
    }
    int ok = 1;
    if (last) {
      __hip_atomic_store(&s->round.v, target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      const unsigned long long t0 = __builtin_amdgcn_s_memrealtime();
      unsigned int polls = 0;
      while (__hip_atomic_load(&s->round.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if ((polls++ & 31u) == 0) {
          if (__hip_atomic_load(&s->abort.v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            ok = 0;
            break;
          }
          if (__builtin_amdgcn_s_memrealtime() - t0 > limit) {
            __hip_atomic_store(&s->abort.v, 1ULL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = 0;
            break;
          }
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (Fenced) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    go = ok;
  }
  __syncthreads();
  return go != 0;
}

}  // namespace dyng_hip
