// rz_redzone_dev.hpp -- device bodies of the three Redzone steps, shared by the
// stream path (one kernel per step), the cooperative single launch and the
// persistent work-queue kernel.
//
//   mark_body   tail bitmap marks (relaxed atomicOr), redzone.hxx:698-707
//   move_slice  one slice of pairs: moves + in-slice orphans, redzone.hxx:714-822
//   pair_body   prefix of per-slice orphan counts + orphan pairing, redzone.hxx:834-934
//
// COH selects coherent (relaxed agent-scope atomic) access for every value that
// another workgroup of the same kernel reads; the stream path relies on the
// kernel boundary instead and uses plain accesses.
#pragma once
#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

// ---------------------------------------------------------------------------
// coherent element access
// ---------------------------------------------------------------------------
template <typename T> __device__ __forceinline__ void st_coh(T* p, T v);
template <> __device__ __forceinline__ void st_coh<uint32_t>(uint32_t* p, uint32_t v) { st_relaxed(p, v); }
template <> __device__ __forceinline__ void st_coh<uint64_t>(uint64_t* p, uint64_t v) { st_relaxed(p, v); }
template <> __device__ __forceinline__ void st_coh<uint4>(uint4* p, uint4 v) {
  uint64_t* q = (uint64_t*)p;
  st_relaxed(q, (uint64_t)v.x | ((uint64_t)v.y << 32));
  st_relaxed(q + 1, (uint64_t)v.z | ((uint64_t)v.w << 32));
}
template <typename T> __device__ __forceinline__ T ld_coh(const T* p);
template <> __device__ __forceinline__ uint32_t ld_coh<uint32_t>(const uint32_t* p) { return ld_relaxed(p); }
template <> __device__ __forceinline__ uint64_t ld_coh<uint64_t>(const uint64_t* p) { return ld_relaxed(p); }
template <> __device__ __forceinline__ uint4 ld_coh<uint4>(const uint4* p) {
  const uint64_t* q = (const uint64_t*)p;
  const uint64_t lo = ld_relaxed(q), hi = ld_relaxed(q + 1);
  return make_uint4((uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32));
}
template <bool COH, typename T> __device__ __forceinline__ void st_x(T* p, T v) {
  if constexpr (COH) st_coh<T>(p, v); else *p = v;
}
template <bool COH, typename T> __device__ __forceinline__ T ld_x(const T* p) {
  if constexpr (COH) return ld_coh<T>(p); else return *p;
}

// ---------------------------------------------------------------------------
// 1. mark: R[0, count) (a window of the removal list of a list with k
// removals and red zone start sZ); thread gtid of `stride` threads.
// ---------------------------------------------------------------------------
// U: removals per thread whose loads are in flight together.
template <bool STATS, int U = 4>
__device__ __forceinline__ void mark_body(const int32_t* __restrict__ R, uint32_t count, uint32_t k, uint32_t sZ,
                                          uint32_t* __restrict__ bits, uint64_t* __restrict__ stat_tail,
                                          uint32_t gtid, uint32_t stride) {
  uint32_t hits = 0;
  for (uint32_t i0 = gtid; i0 < count; i0 += stride * U) {
    uint32_t r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * stride;
      r[u] = i < count ? (uint32_t)R[i] : sZ - 1u;  // sZ-1 is never inside the tail (k > 0 => sZ < n)
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t z = r[u] - sZ;  // wraps to >= 2^31 > k below the tail
      if (z < k) {
        or_relaxed(&bits[z >> 5], 1u << (z & 31));
        if (STATS) ++hits;
      }
    }
  }
  if (STATS) {
    hits = wave_sum(hits);
    if (lane_id() == 0 && hits) add_relaxed(stat_tail, (uint64_t)hits);
  }
}

// ---------------------------------------------------------------------------
// 2. move and collect orphans
// ---------------------------------------------------------------------------
// Item classes.  A pair with a dead source and an invalid destination does nothing.
enum : uint32_t { kNone = 0, kMove = 1, kSrcOrphan = 2, kDstOrphan = 3 };

template <int WG, int IPT>
struct MoveShared {
  static constexpr int E = IPT * (WG / kWave);
  static constexpr int WORDS = WG * IPT / 32;  // bitmap words of one step
  uint32_t off[E + 1];
  uint32_t bits[WORDS];
};

// One slice [a, b) of pairs, a = slice * L.  If `signal` is not null it is
// incremented (relaxed) once the slice's orphans and counts are visible.
template <typename T, int WG, int IPT, bool COH>
__device__ __forceinline__ void move_slice(T* __restrict__ A, const int32_t* __restrict__ R, int32_t* __restrict__ Rd,
                                           uint32_t k, uint32_t sZ, uint32_t L, uint32_t slice,
                                           uint32_t* __restrict__ bits, uint64_t* __restrict__ counts,
                                           MoveShared<WG, IPT>& sh, uint32_t* __restrict__ signal = nullptr) {
  constexpr int NW = WG / kWave;
  constexpr int E = MoveShared<WG, IPT>::E;  // (step item, wave) count cells, in i order
  constexpr int PER = (E + kWave - 1) / kWave;
  uint32_t* s_off = sh.off;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint32_t a = slice * L;
  const uint32_t b = min(a + L, k);
  T* __restrict__ Z = A + sZ;
  uint32_t carry_s = 0, carry_d = 0;

  for (uint32_t base = a; base < b; base += WG * IPT) {
    uint32_t d[IPT], wv[IPT], info[IPT];
    T z[IPT];
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t p = base + j * WG + tid;
      if (p < b) {
        d[j] = (uint32_t)R[p];
        z[j] = Z[p];
        if constexpr (!COH) wv[j] = bits[p >> 5];
      } else {
        d[j] = sZ;  // invalid destination
      }
    }
    if constexpr (COH) {
      // Coherent loads bypass the caches: take each bitmap word of the step
      // once (base is word aligned) and share it through LDS.
      constexpr int WORDS = MoveShared<WG, IPT>::WORDS;
      const uint32_t wend = (b + 31u) >> 5;
      for (int w = tid; w < WORDS; w += WG) {
        const uint32_t gw = (base >> 5) + w;
        sh.bits[w] = gw < wend ? ld_relaxed(&bits[gw]) : 0xffffffffu;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < IPT; ++j) wv[j] = sh.bits[(j * WG + tid) >> 5];
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      if (base + j * WG + tid >= b) wv[j] = 0xffffffffu;  // dead source
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t p = base + j * WG + tid;
      const bool dead = (wv[j] >> (p & 31)) & 1u;
      const bool dst_ok = d[j] < sZ;
      const uint32_t cls = dead ? (dst_ok ? kDstOrphan : kNone) : (dst_ok ? kMove : kSrcOrphan);
      const uint64_t ms = ballot(cls == kSrcOrphan);
      const uint64_t md = ballot(cls == kDstOrphan);
      info[j] = cls | (lane_rank(ms) << 8) | (lane_rank(md) << 16);
      if (lane == 0) s_off[j * NW + wave] = (uint32_t)__popcll(ms) | ((uint32_t)__popcll(md) << 16);
    }
    // Every Z and R load of this step must have returned before any thread
    // writes an orphan into this slice's own Z / R range.
    wait_all_memory();
    __syncthreads();
    if (wave == 0) {
      uint32_t loc[PER];
      uint32_t sum = 0;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int idx = lane * PER + q;
        loc[q] = idx < E ? s_off[idx] : 0u;
        sum += loc[q];
      }
      const uint32_t x = wave_incl_scan(sum);
      uint32_t ex = x - sum;
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int idx = lane * PER + q;
        if (idx < E) s_off[idx] = ex;
        ex += loc[q];
      }
      if (lane == kWave - 1) s_off[E] = x;
    }
    __syncthreads();
    // orphans first ...
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t cls = info[j] & 0xffu;
      if (cls != kSrcOrphan && cls != kDstOrphan) continue;
      const uint32_t off = s_off[j * NW + wave];
      if (cls == kSrcOrphan) {
        st_x<COH>(&Z[a + carry_s + (off & 0xffffu) + ((info[j] >> 8) & 0xffu)], z[j]);
      } else {
        st_x<COH>((uint32_t*)&Rd[a + carry_d + (off >> 16) + (info[j] >> 16)], d[j]);
      }
    }
    const uint32_t tot = s_off[E];
    carry_s += tot & 0xffffu;
    carry_d += tot >> 16;
    if (base + WG * IPT >= b) {
      // ... then the slice's counts (and the optional signal that they and
      // the orphans are visible), before the scattered moves are issued, so
      // that the signal does not wait for the moves to drain
      if (tid == 0) st_x<COH>(&counts[slice], (uint64_t)carry_d | ((uint64_t)carry_s << 32));
      if (signal) {
        wait_all_memory();
        __syncthreads();
        if (tid == 0) add_relaxed(signal, 1u);
      }
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if ((info[j] & 0xffu) == kMove) A[d[j]] = z[j];
    __syncthreads();
  }
  // leave the bitmap zero for the next call (slices are word aligned)
  for (uint32_t w = (a >> 5) + tid; w < ((b + 31u) >> 5); w += WG) bits[w] = 0u;
}

// ---------------------------------------------------------------------------
// 3. pair orphans
// ---------------------------------------------------------------------------
template <int WG, int MAXS>
struct PairShared {
  uint64_t pre[MAXS + 1];
  uint64_t tmp[WG / kWave + 1];
};

// Exclusive prefix of the per-slice counts into LDS (s_pre: ns+1 entries,
// s_tmp: WG/64+1 entries), then pairs t = block*WG + tid, stepping by
// nblocks*WG.  Dest orphans are the low 32 bits of a count word.
template <typename T, int WG, bool COH>
__device__ __forceinline__ void pair_body_lds(T* __restrict__ A, const int32_t* __restrict__ Rd, uint32_t k,
                                              uint32_t sZ, uint32_t L, uint32_t ns, const uint64_t* __restrict__ counts,
                                              uint32_t* __restrict__ ctrl, int stats, uint32_t block,
                                              uint32_t nblocks, uint64_t* s_pre, uint64_t* s_tmp) {
  const int tid = threadIdx.x;
  // exclusive prefix in chunks of 4*WG counts (4 consecutive counts per thread)
  uint64_t carry = 0;
  for (uint32_t c0 = 0; c0 < ns; c0 += WG * 4) {
    uint64_t loc[4];
    uint64_t sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t s = c0 + tid * 4 + q;
      loc[q] = s < ns ? ld_x<COH>(&counts[s]) : 0ull;
      sum += loc[q];
    }
    uint64_t chunk_total;
    uint64_t ex = carry + block_excl_scan<WG>(sum, s_tmp, &chunk_total);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t s = c0 + tid * 4 + q;
      if (s < ns) s_pre[s] = ex;
      ex += loc[q];
    }
    carry += chunk_total;
  }
  const uint64_t total = carry;
  if (tid == 0) s_pre[ns] = total;
  __syncthreads();
  const uint32_t tot_d = (uint32_t)total, tot_s = (uint32_t)(total >> 32);
  const uint32_t pairs = min(tot_d, tot_s);
  if (block == 0 && tid == 0) {
    if (tot_d != tot_s) or_relaxed(&ctrl[kCtrlError], kErrOrphanMismatch);
    if (stats) st_relaxed((uint64_t*)&ctrl[kCtrlStatOrphans], (uint64_t)pairs);
  }
  T* __restrict__ Z = A + sZ;
  for (uint32_t t = block * WG + tid; t < pairs; t += nblocks * WG) {
    const int sd = upper_row((int)ns, t, [&](int s) { return (uint32_t)s_pre[s]; });
    const int ss = upper_row((int)ns, t, [&](int s) { return (uint32_t)(s_pre[s] >> 32); });
    // With valid counts both offsets lie inside their slice; the guards keep a
    // wait that gave up (stale counts) from indexing outside R, Z or the head.
    const uint32_t od = t - (uint32_t)s_pre[sd], os = t - (uint32_t)(s_pre[ss] >> 32);
    if (od >= L || os >= L) continue;
    const uint32_t id = (uint32_t)sd * L + od, is = (uint32_t)ss * L + os;
    if (id >= k || is >= k) continue;
    const uint32_t dst = ld_x<COH>((const uint32_t*)&Rd[id]);
    if (dst < sZ) A[dst] = ld_x<COH>(&Z[is]);
  }
  __syncthreads();  // s_pre may be reused by the caller
}

template <typename T, int WG, int MAXS, bool COH>
__device__ __forceinline__ void pair_body(T* __restrict__ A, const int32_t* __restrict__ Rd, uint32_t k, uint32_t sZ,
                                          uint32_t L, uint32_t ns, const uint64_t* __restrict__ counts,
                                          uint32_t* __restrict__ ctrl, int stats, uint32_t block,
                                          uint32_t nblocks, PairShared<WG, MAXS>& sh) {
  pair_body_lds<T, WG, COH>(A, Rd, k, sZ, L, ns, counts, ctrl, stats, block, nblocks, sh.pre, sh.tmp);
}

// ---------------------------------------------------------------------------
// in-kernel dependencies between workgroups (relaxed counters)
// ---------------------------------------------------------------------------
// Everything this workgroup wrote is visible, then the counter goes up by one.
__device__ __forceinline__ void signal_done(uint32_t* counter) {
  wait_all_memory();
  __syncthreads();
  if (threadIdx.x == 0) add_relaxed(counter, 1u);
}

// Wait (thread 0 polls, the workgroup waits at a barrier) until the counter
// reaches target; a bounded wait that flags kErrBarrierTimeout if it gives up
// (after spin_cap polls; the stall hook makes every poll fail).
__device__ __forceinline__ void wait_count(const uint32_t* counter, uint32_t target, uint32_t* ctrl) {
  if (threadIdx.x == 0) {
    uint32_t spins = 0;
    const uint32_t cap = spin_cap(ctrl, 1u << 22);
    const bool stall = stalled(ctrl);
    while (stall || ld_relaxed(counter) < target) {
      __builtin_amdgcn_s_sleep(1);
      if (++spins > cap) {
        or_relaxed(&ctrl[kCtrlError], kErrBarrierTimeout);
        break;
      }
    }
  }
  __syncthreads();
}

// Wait (thread 0 polls, the workgroup waits at a barrier) until *word ==
// value, with a pause that doubles from 64 to 64*cap cycles between polls; a
// bounded wait that flags kErrBarrierTimeout if it gives up.
__device__ __forceinline__ void wait_value_backoff(const uint32_t* word, uint32_t value, uint32_t* ctrl,
                                                   uint32_t cap) {
  if (threadIdx.x == 0) {
    uint32_t spins = 0, pause = 1;
    const uint32_t limit = spin_cap(ctrl, 1u << 22);
    const bool stall = stalled(ctrl);
    while (stall || ld_relaxed(word) != value) {
      for (uint32_t i = 0; i < pause; ++i) __builtin_amdgcn_s_sleep(1);
      if (pause < cap) pause <<= 1;
      if (++spins > limit) {
        or_relaxed(&ctrl[kCtrlError], kErrBarrierTimeout);
        break;
      }
    }
  }
  __syncthreads();
}

}  // namespace rz
