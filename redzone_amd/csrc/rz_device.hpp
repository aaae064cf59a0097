// rz_device.hpp -- wave64 / LDS / coherence helpers shared by the gfx950 kernels.
//
// Cross-workgroup communication inside one kernel never uses acquire/release
// fences at agent scope: on MI355X each of those writes back or invalidates a
// whole XCD L2 and costs microseconds per workgroup.  Instead every word that
// another workgroup reads inside the same kernel is written and read with
// relaxed agent-scope atomics (per-location coherent across XCDs), and the
// ordering that matters is carried by data dependencies or by an explicit
// wait for outstanding memory operations.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rz {

constexpr int kWave = 64;

template <int E> struct elem_of;
template <> struct elem_of<4> { using type = uint32_t; };
template <> struct elem_of<8> { using type = uint64_t; };
template <> struct elem_of<16> { using type = uint4; };

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// number of set bits of `m` below this lane (the lane's rank among the set lanes)
__device__ __forceinline__ uint32_t lane_rank(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// LDS counter increment with a rank: every active lane adds one to cnt[bin]
// and gets a rank that is unique within the bin.  When all active lanes of
// the wave share one bin (a run of removals), one atomic adds them all; else
// each lane takes its own (grouping partly shared bins costs more on uniform
// lists than it saves on clustered ones).  Call from every lane of the wave.
__device__ __forceinline__ uint32_t lds_bin_add(uint32_t* cnt, uint32_t bin, bool active) {
  const uint64_t act = ballot(active);
  if (!act) return 0u;
  const int leader = __builtin_ctzll(act);
  const uint32_t lb = (uint32_t)__builtin_amdgcn_readlane((int)bin, leader);
  if (ballot(active && bin == lb) == act) {
    uint32_t base = 0u;
    if (lane_id() == leader) base = atomicAdd(&cnt[lb], (uint32_t)__popcll(act));
    return (uint32_t)__builtin_amdgcn_readlane((int)base, leader) + lane_rank(act);
  }
  return active ? atomicAdd(&cnt[bin], 1u) : 0u;
}
// Count only (no ranks), grouped the same way.
__device__ __forceinline__ void lds_bin_count(uint32_t* cnt, uint32_t bin, bool active) {
  const uint64_t act = ballot(active);
  if (!act) return;
  const int leader = __builtin_ctzll(act);
  const uint32_t lb = (uint32_t)__builtin_amdgcn_readlane((int)bin, leader);
  if (ballot(active && bin == lb) == act) {
    if (lane_id() == leader) atomicAdd(&cnt[lb], (uint32_t)__popcll(act));
  } else if (active) {
    atomicAdd(&cnt[bin], 1u);
  }
}

__device__ __forceinline__ uint32_t ld_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_relaxed(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(uint64_t* p, uint64_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t or_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t and_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_and(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t add_relaxed(uint64_t* p, uint64_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t add_relaxed(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t xchg_relaxed(uint64_t* p, uint64_t v) {
  return __hip_atomic_exchange(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wait until every memory operation this wave issued has completed (gfx9
// s_waitcnt with all counters at zero; on gfx950 vmcnt also counts stores).
// The signal fences keep the compiler from moving memory accesses across it.
__device__ __forceinline__ void wait_all_memory() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Inclusive scan across the 64 lanes of a wave.
template <typename V>
__device__ __forceinline__ V wave_incl_scan(V v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    V t = __shfl_up(v, o, kWave);
    if (lane >= o) v += t;
  }
  return v;
}

template <typename V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Exclusive scan of one value per thread across a workgroup of WG threads.
// s_tmp needs WG/64 + 1 entries.  Returns the exclusive prefix; *total gets the sum.
template <int WG, typename V>
__device__ __forceinline__ V block_excl_scan(V v, V* s_tmp, V* total) {
  constexpr int NW = WG / kWave;
  const int lane = lane_id(), wave = threadIdx.x / kWave;
  V x = wave_incl_scan(v);
  if (lane == kWave - 1) s_tmp[wave] = x;
  __syncthreads();
  if (wave == 0) {
    V w = lane < NW ? s_tmp[lane] : V(0);
    V wx = wave_incl_scan(w);
    if (lane < NW) s_tmp[lane] = wx - w;  // exclusive wave offsets
    if (lane == NW - 1) s_tmp[NW] = wx;
  }
  __syncthreads();
  V r = s_tmp[wave] + x - v;
  *total = s_tmp[NW];
  __syncthreads();
  return r;
}

// Largest s in [0, cnt) with key(pre[s]) <= t, for a non-decreasing prefix.
template <typename F>
__device__ __forceinline__ int upper_row(int cnt, uint32_t t, F key) {
  int lo = 0, hi = cnt - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (key(mid) <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace rz
