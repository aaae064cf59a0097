// rz_compact_dev.hpp -- one tile of the O(n) stable in-place compaction:
// the standalone partition's rz_compact (rz_partition.hip, one tile per
// workgroup) and the batch kernel's partition items (rz_batch.hip, one tile per
// queue item).
//
// The tile ranks its kept items with wave ballots and one LDS scan,
// publishes its count, finds its output offset by decoupled look-back over the
// earlier tiles' status words, and writes its kept items in order.  A tile
// publishes only after all of its loads have returned, and writes only below
// the end of its own input, so the in-place pass never overwrites an item
// before it was read.  It clears the bitmap words it read.
#pragma once

#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

// status word: [63:62] flag (0 none, 1 aggregate, 2 inclusive prefix), [61:32] epoch, [31:0] count
constexpr uint64_t kFlagAgg = 1ull, kFlagIncl = 2ull;
__device__ __forceinline__ uint64_t pack_status(uint64_t flag, uint32_t epoch, uint32_t v) {
  return (flag << 62) | ((uint64_t)(epoch & 0x3fffffffu) << 32) | v;
}

// Tile t of an n-item list whose removals are marked in `bits`.  Status words
// count as published when their epoch field equals `epoch` (a caller that
// zeroes them before the pass passes epoch 0).  COH: read the bitmap with
// coherent loads (it was marked in the same launch).
template <typename T, int WG, int IPT, bool COH>
__device__ __forceinline__ void compact_tile(T* __restrict__ A, uint32_t n, uint32_t* __restrict__ bits,
                                             uint64_t* __restrict__ status, uint32_t t, uint32_t epoch,
                                             uint32_t* __restrict__ ctrl) {
  constexpr int NW = WG / kWave;
  constexpr int E = IPT * NW;
  constexpr int PER = (E + kWave - 1) / kWave;
  __shared__ uint32_t s_off[E + 1];
  __shared__ uint32_t s_excl;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint32_t base = t * (uint32_t)(WG * IPT);

  T x[IPT];
  uint32_t rank[IPT];
  uint32_t keepbits = 0;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t p = base + j * WG + tid;
    uint32_t w = 0xffffffffu;
    x[j] = T();  // defined in every lane: keeps 16-byte items in registers
    if (p < n) {
      x[j] = A[p];
      w = COH ? ld_relaxed(&bits[p >> 5]) : bits[p >> 5];
    }
    const bool keep = !((w >> (p & 31)) & 1u);
    const uint64_t m = ballot(keep);
    rank[j] = lane_rank(m);
    keepbits |= (uint32_t)keep << j;
    if (lane == 0) s_off[j * NW + wave] = (uint32_t)__popcll(m);
  }
  wait_all_memory();  // all item loads of this tile have returned
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
    const uint32_t xs = wave_incl_scan(sum);
    uint32_t ex = xs - sum;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int idx = lane * PER + q;
      if (idx < E) s_off[idx] = ex;
      ex += loc[q];
    }
    const uint32_t agg = __shfl(xs, kWave - 1, kWave);
    // decoupled look-back
    uint32_t excl = 0;
    if (t == 0) {
      if (lane == 0) st_relaxed(&status[0], pack_status(kFlagIncl, epoch, agg));
    } else {
      if (lane == 0) st_relaxed(&status[t], pack_status(kFlagAgg, epoch, agg));
      int64_t j = (int64_t)t - 1;
      uint32_t spins = 0;
      while (true) {
        const int64_t q = j - lane;
        uint64_t st = q >= 0 ? ld_relaxed(&status[q]) : pack_status(kFlagIncl, epoch, 0u);
        uint32_t flag = (((uint32_t)(st >> 32) & 0x3fffffffu) == epoch) ? (uint32_t)(st >> 62) : 0u;
        if (q < 0) flag = (uint32_t)kFlagIncl;
        const uint64_t incl_m = ballot(flag == kFlagIncl);
        const uint64_t none_m = ballot(flag == 0u);
        // lanes up to and including the nearest inclusive prefix must all be ready
        const uint64_t upto = incl_m ? ((incl_m & (~incl_m + 1)) << 1) - 1 : ~0ull;
        if (none_m & upto) {
          if (++spins > spin_cap(ctrl, 1u << 24)) {  // give up (the grid is not making progress)
            if (lane == 0) or_relaxed(&ctrl[kCtrlError], kErrLookbackTimeout);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        const uint32_t v = ((upto >> lane) & 1ull) ? (uint32_t)st : 0u;
        excl += wave_sum(v);
        if (incl_m) break;
        j -= kWave;
      }
      if (lane == 0) st_relaxed(&status[t], pack_status(kFlagIncl, epoch, excl + agg));
    }
    if (lane == 0) s_excl = excl;
  }
  __syncthreads();
  const uint32_t excl = s_excl;
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    if ((keepbits >> j) & 1u) {
      const uint32_t pos = excl + s_off[j * NW + wave] + rank[j];
      if (pos < n) A[pos] = x[j];
    }
  }
  // clear the bitmap words this tile read (tiles are word aligned)
  const uint32_t w0 = base >> 5, w1 = min((base + WG * IPT + 31u) >> 5, (n + 31u) >> 5);
  for (uint32_t w = w0 + tid; w < w1; w += WG) bits[w] = 0u;
}

}  // namespace rz
