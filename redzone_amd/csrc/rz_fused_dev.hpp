// rz_fused_dev.hpp -- the fused move + pair body: one slice moves its pairs,
// keeps its orphans in LDS, learns the earlier slices' orphan counts by
// decoupled look-back and meets its orphans' partners through exchange words.
// Used by the single-list fused forms (rz_plain.hip: rz_move_pair,
// rz_mark_move_pair, rz_mark_move_pair_sized) and by the batch kernel's fused
// items (rz_batch.hip).
//
// The move step writes no orphans to R or Z.  A slice keeps its orphans in
// LDS (dest positions from the front of an L-entry list, source tail indices
// from the back; a pair yields at most one orphan, so L entries suffice),
// publishes its two orphan counts, and learns the counts of all earlier
// slices by decoupled look-back over per-slice status words.  Slices are
// dispatched (or dequeued) in id order and never wait on a later slice, so the
// look-back always finishes.  Orphan rank t then meets its partner through one
// exchange word per rank: each side swaps in (tag << 32 | own position + 1);
// the side that gets back its partner's entry of the same tag performs the
// move A[dest] = Z[src] and clears the word.  The t-th dest orphan still pairs
// with the t-th source orphan in i order (redzone.hxx:834-934), so the array
// is the same as with the three-kernel form.  Nothing in R and nothing in the
// red zone is written, so RZ_PRESERVE_REMOVALS costs nothing here, and no
// step waits for its predecessor's scattered stores to drain.
//
// Status word: [63:62] flag (0 none, 1 aggregate, 2 inclusive prefix),
// [61:31] dest orphans, [30:0] source orphans (each count is below 2^31, so
// packed words add without carries between the fields).  The caller hands in
// ns status words that are zero, and exchange words whose tag differs from
// `tag` (or that are zero).
#pragma once

#include "rz_redzone_dev.hpp"

namespace rz {

constexpr uint64_t kStAgg = 1ull << 62, kStIncl = 2ull << 62, kStCount = (1ull << 62) - 1;

// COH: read the tail bitmap with coherent loads (it was marked in the same
// launch).  READY: after the first step's loads of R and the red zone, wait
// until the mark workgroups of this launch have raised the ready word (tag + 1)
// of this XCD.  `my_stamps`: this workgroup's row of RZ_PROFILE stage stamps
// (or null).
template <typename T, int WG, int IPT, bool COH, bool READY>
__device__ __forceinline__ void move_pair_core(T* __restrict__ A, const int32_t* __restrict__ R, uint32_t k,
                                               uint32_t sZ, uint32_t L, uint32_t ns, uint32_t* __restrict__ bits,
                                               uint64_t* __restrict__ st, uint64_t* __restrict__ slots,
                                               uint32_t slot_cap, uint32_t* __restrict__ ctrl, int stats, int early,
                                               uint32_t slice, uint32_t backoff, uint32_t tag,
                                               uint64_t* __restrict__ my_stamps) {
  constexpr int NW = WG / kWave;
  constexpr int E = MoveShared<WG, IPT>::E;
  constexpr int PER = (E + kWave - 1) / kWave;
  __shared__ MoveShared<WG, IPT> sh;
  __shared__ uint64_t s_excl;
  extern __shared__ uint32_t s_orph[];  // L entries
  uint32_t* s_off = sh.off;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint32_t a = slice * L, b = min(a + L, k);
  T* __restrict__ Z = A + sZ;

  // one step's loads: R and the red zone, and the bitmap words (coherent
  // loads are made after the step's other loads, below)
  auto load_rz = [&](uint32_t base_, uint32_t (&dd)[IPT], T (&zz)[IPT], uint32_t (&ww)[IPT]) {
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t p = base_ + j * WG + tid;
      if (p < b) {
        dd[j] = (uint32_t)R[p];
        zz[j] = Z[p];
        ww[j] = COH ? 0u : bits[p >> 5];
      } else {
        dd[j] = sZ;           // invalid destination
        zz[j] = T();          // (a defined value keeps 16-byte items in registers)
        ww[j] = 0xffffffffu;  // dead source
      }
    }
  };
  uint32_t carry_s = 0, carry_d = 0;
  // Exclusive orphan counts of the earlier slices, by decoupled look-back,
  // into s_excl (wave 0).  The slice's own aggregate is published first, and
  // a slice never waits on a later one.
  auto lookback = [&]() {
    const uint32_t mm = early ? min(carry_d, carry_s) : 0u;
    const uint64_t agg = ((uint64_t)(carry_d - mm) << 31) | (uint64_t)(carry_s - mm);
    uint64_t excl = 0;
    if (slice > 0) {
      int64_t j = (int64_t)slice - 1;
      uint32_t spins = 0;
      const uint32_t cap = spin_cap(ctrl, 1u << 24);
      const bool stall = stalled(ctrl);
      while (true) {
        const int64_t q = j - lane;
        const uint64_t w = q >= 0 ? ld_relaxed(&st[q]) : kStIncl;
        const uint32_t flag = (uint32_t)(w >> 62);
        const uint64_t incl_m = ballot(flag == 2u);
        const uint64_t none_m = ballot(flag == 0u);
        // lanes up to and including the nearest inclusive prefix must all be ready
        const uint64_t upto = incl_m ? ((incl_m & (~incl_m + 1)) << 1) - 1 : ~0ull;
        if (stall || (none_m & upto)) {
          // a give-up leaves excl short: the ranks below stay inside the
          // exchange words and every move below checks its indices
          if (++spins > cap) {
            if (lane == 0) or_relaxed(&ctrl[kCtrlError], kErrLookbackTimeout);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        excl += wave_sum(((upto >> lane) & 1ull) ? (w & kStCount) : 0ull);
        if (incl_m) break;
        j -= kWave;
      }
      if (lane == 0) st_relaxed(&st[slice], kStIncl | (excl + agg));
    }
    if (lane == 0) {
      s_excl = excl;
      if (slice == ns - 1) {
        const uint64_t tot = excl + agg;
        const uint32_t td = (uint32_t)(tot >> 31), ts = (uint32_t)tot & 0x7fffffffu;
        if (td != ts) or_relaxed(&ctrl[kCtrlError], kErrOrphanMismatch);
        if (stats) st_relaxed((uint64_t*)&ctrl[kCtrlStatOrphans], (uint64_t)min(td, ts));
      }
    }
  };
  for (uint32_t base = a; base < b; base += WG * IPT) {
    uint32_t d[IPT], wv[IPT], info[IPT];
    T z[IPT];
    load_rz(base, d, z, wv);
    if constexpr (READY) {
      // every mark of this launch is in: this XCD's copy of the ready word
      if (base == a) {
        wait_value_backoff(&ctrl[kCtrlMarkReady + 32 * (blockIdx.x & 7u)], tag + 1u, ctrl, backoff);
        if (my_stamps && tid == 0) my_stamps[kStampReady] = wall_clock64();
      }
    }
    if constexpr (COH) {
#pragma unroll
      for (int j = 0; j < IPT; ++j) {
        const uint32_t p = base + j * WG + tid;
        if (p < b) wv[j] = ld_relaxed(&bits[p >> 5]);
      }
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
      if (lane == kWave - 1) {
        s_off[E] = x;
        if (base + WG * IPT >= b) {
          // the slice's counts are final: publish them before the last moves
          // (with early pairing, only the orphans left after the slice pairs its own)
          const uint32_t td = carry_d + (x >> 16), ts = carry_s + (x & 0xffffu);
          const uint32_t m = early ? min(td, ts) : 0u;
          const uint64_t agg = ((uint64_t)(td - m) << 31) | (uint64_t)(ts - m);
          st_relaxed(&st[slice], (slice == 0 ? kStIncl : kStAgg) | agg);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t cls = info[j] & 0xffu;
      const uint32_t off = s_off[j * NW + wave];
      if (cls == kDstOrphan)
        s_orph[carry_d + (off >> 16) + (info[j] >> 16)] = d[j];
      else if (cls == kSrcOrphan)
        s_orph[L - 1u - (carry_s + (off & 0xffffu) + ((info[j] >> 8) & 0xffu))] = base + j * WG + tid;
    }
    const uint32_t tot = s_off[E];
    carry_s += tot & 0xffffu;
    carry_d += tot >> 16;
    if (wave == 0 && base + WG * IPT >= b) {
      // The last step: wave 0 looks back before it issues its own moves.  A
      // wave's loads wait (vmcnt) for its earlier scattered stores to drain,
      // so a look-back after the moves would start only once they had; the
      // other waves' moves drain meanwhile.
      lookback();
      if (my_stamps && tid == 0) my_stamps[kStampLookback] = wall_clock64();
    }
#pragma unroll
    for (int j = 0; j < IPT; ++j)
      if ((info[j] & 0xffu) == kMove) A[d[j]] = z[j];
    __syncthreads();
  }

  if (my_stamps && tid == 0) my_stamps[kStampMoves] = wall_clock64();
  // Early pairing (RZ_ANY_PAIRING): the slice's first m dest orphans take its
  // first m source orphans here, and only the rest go to the global exchange.
  const uint32_t m = early ? min(carry_d, carry_s) : 0u;
  for (uint32_t j = tid; j < m; j += WG) {
    const uint32_t dst = s_orph[j], src = s_orph[L - 1u - j];
    if (dst < sZ && src < k) A[dst] = Z[src];
  }
  const uint32_t rd = carry_d - m, rs = carry_s - m;
  __syncthreads();  // the early pairs have read their LDS entries

  // meet the partners
  const uint64_t excl = s_excl;
  const uint32_t xd = (uint32_t)(excl >> 31), xs = (uint32_t)excl & 0x7fffffffu;
  const uint64_t tagged = (uint64_t)tag << 32;
  const uint32_t norph = rd + rs;
  for (uint32_t e = tid; e < norph; e += WG) {
    const bool is_d = e < rd;
    const uint32_t mine = is_d ? s_orph[m + e] : s_orph[L - 1u - (m + e - rd)];
    const uint32_t t = is_d ? xd + e : xs + (e - rd);
    if (t >= slot_cap) continue;  // only with an invalid removal list
    const uint64_t old = xchg_relaxed(&slots[t], tagged | (uint64_t)(mine + 1u));
    if ((uint32_t)(old >> 32) != tag || (uint32_t)old == 0u) continue;  // first to arrive
    const uint32_t other = (uint32_t)old - 1u;
    if (is_d) {
      if (other < k) A[mine] = Z[other];
    } else {
      if (other < sZ) A[other] = Z[mine];
    }
    st_relaxed(&slots[t], 0ull);
  }
  if (my_stamps) {
    __syncthreads();
    if (tid == 0) my_stamps[kStampExchange] = wall_clock64();
  }
  // leave the bitmap zero for the next call (slices are word aligned)
  for (uint32_t w = (a >> 5) + tid; w < ((b + 31u) >> 5); w += WG) bits[w] = 0u;
}

}  // namespace rz
