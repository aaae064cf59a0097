// rz_plain.hip -- Redzone O(k) removal from one list on gfx950.
//
// The removal list R (k entries) pairs with the red zone Z = A[n-k, n): pair i
// moves Z[i] into hole R[i].  Three dependent steps (bodies in
// rz_redzone_dev.hpp):
//
//   1. mark   every R[i] >= n-k kills tail item R[i]-(n-k): one bit in a k-bit
//             tail bitmap (relaxed atomicOr).  Replaces the sentinel write of
//             redzone.hxx:698-707, so no payload value is reserved.
//   2. move   R is cut into slices of L pairs, one workgroup each.  A
//             workgroup loads its R, Z and bitmap words coalesced, classifies
//             every pair with two 64-lane ballots, and ranks the orphans of its
//             slice in i order with one LDS scan per step (redzone.hxx:714-822
//             does two block scans per step on 32-lane warps).  Moves go
//             straight to their holes; dest orphans go to the slice's own part
//             of R, source orphans to the slice's own part of Z.  The slice's
//             bitmap words are cleared after use, so the bitmap is zero again
//             for the next call without a memset.
//   3. pair   every workgroup scans the per-slice orphan counts in LDS and
//             pairs the t-th dest orphan with the t-th source orphan
//             (redzone.hxx:834-934).  Orphans are ordered by slice then by i,
//             so the result is the canonical one for every launch shape.
//
// Launch variants: (0) one kernel per step on the caller's stream, the kernel
// boundary being the grid-wide barrier; (1) one cooperative kernel with
// in-kernel barriers made of relaxed atomics (RemoveFromList_GPUWide's
// structure, redzone.hxx:626-941); (2) the persistent work-queue kernel of
// rz_batch.hip with a single list; (3) one ordered launch, roles by workgroup
// id; (4) mark, then one kernel that moves and pairs orphans through a
// decoupled look-back and per-rank exchange words (rz_move_pair).
//
// Contract violations (duplicates, out-of-range indices) never cause an
// out-of-bounds access: marks are taken only for indices inside the tail,
// moves only into holes below n-k, and pairing stops at the smaller orphan
// total (a mismatch sets an error bit in the control block).
#include "../../include/redzone_device.h"
#include "rz_fused_dev.hpp"
#include "rz_redzone_dev.hpp"

namespace rz {

// The mark step for callers that mark into their own bitmap (rz_mark_removals),
// written with the public producer helper.
__global__ __launch_bounds__(256) void rz_mark_removals_kernel(const int32_t* __restrict__ R, int64_t n, int64_t k,
                                                                uint32_t* __restrict__ bits) {
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < k; i += (int64_t)gridDim.x * 256ll)
    rz_mark_removal(bits, R[i], n, k);
}

rz_status launch_mark_removals(rz_ctx* ctx, const int32_t* R, int64_t n, int64_t k, uint32_t* bits, hipStream_t s) {
  if (k == 0) return RZ_OK;
  const int64_t g = (k + 1023) / 1024;
  const int64_t cap = (int64_t)ctx->cu_count * 8;
  hipLaunchKernelGGL(rz_mark_removals_kernel, dim3((uint32_t)(g < cap ? g : cap)), dim3(256), 0, s, R, n, k, bits);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

template <bool STATS>
__global__ __launch_bounds__(256) void rz_mark_tail(const int32_t* __restrict__ R, uint32_t k, uint32_t sZ,
                                                     uint32_t* __restrict__ bits, uint64_t* __restrict__ stat_tail) {
  mark_body<STATS>(R, k, k, sZ, bits, stat_tail, blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u);
}

// Sorted removal list: the removals that fall in the red zone are R's suffix.
// Workgroup w takes the chunk R[w*C, (w+1)*C) and leaves at once when the
// chunk's last entry is below the red zone, so only the suffix is read.
constexpr uint32_t kSortedChunk = 4096;
__global__ __launch_bounds__(256) void rz_mark_tail_sorted(const int32_t* __restrict__ R, uint32_t k, uint32_t sZ,
                                                            uint32_t* __restrict__ bits) {
  const uint32_t lo = blockIdx.x * kSortedChunk;
  const uint32_t hi = min(lo + kSortedChunk, k);
  if ((uint32_t)R[hi - 1] < sZ) return;  // uniform across the workgroup
  mark_body<false>(R + lo, hi - lo, k, sZ, bits, nullptr, threadIdx.x, 256u);
}

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_move(T* __restrict__ A, const int32_t* __restrict__ R,
                                               int32_t* __restrict__ Rd, uint32_t k, uint32_t sZ, uint32_t L,
                                               uint32_t* __restrict__ bits, uint64_t* __restrict__ counts) {
  __shared__ MoveShared<WG, IPT> sh;
  move_slice<T, WG, IPT, false>(A, R, Rd, k, sZ, L, blockIdx.x, bits, counts, sh);
}

template <typename T>
__global__ __launch_bounds__(256) void rz_pair(T* __restrict__ A, const int32_t* __restrict__ Rd, uint32_t k, uint32_t sZ,
                                                uint32_t L, uint32_t ns, const uint64_t* __restrict__ counts,
                                                uint32_t* __restrict__ ctrl, int stats) {
  __shared__ PairShared<256, kMaxSlices> sh;
  pair_body<T, 256, kMaxSlices, false>(A, Rd, k, sZ, L, ns, counts, ctrl, stats, blockIdx.x, gridDim.x, sh);
}

// ---------------------------------------------------------------------------
// fused move + pair (variant 4): mark, then one kernel that moves and pairs
// (body: move_pair_core, rz_fused_dev.hpp).  The status words live in two
// halves of kMaxSlices used on alternate calls (parity of the call epoch,
// which the last workgroup to leave bumps for the next call); a call zeroes
// the half it does not use, so the next call starts from zeros.  The epoch is
// also the exchange words' tag.
// ---------------------------------------------------------------------------
// COH: the bitmap was marked in the same launch (rz_mark_move_pair): after the
// first step's loads of R and the red zone, wait until the mark workgroups
// have raised the ready word (this call's epoch + 1), then read bitmap words
// with coherent loads.
template <typename T, int WG, int IPT, bool COH>
__device__ __forceinline__ void move_pair_body(T* __restrict__ A, const int32_t* __restrict__ R, uint32_t k,
                                               uint32_t sZ, uint32_t L, uint32_t ns, uint32_t* __restrict__ bits,
                                               uint64_t* __restrict__ mstatus, uint64_t* __restrict__ slots,
                                               uint32_t slot_cap, uint32_t* __restrict__ ctrl, int stats,
                                               int early, uint32_t slice, uint32_t backoff,
                                               uint32_t* __restrict__ err_out, uint64_t* __restrict__ stamps) {
  const int tid = threadIdx.x;
  // RZ_PROFILE: thread 0 stamps the wall clock at each stage boundary
  uint64_t* const my_stamps = stamps ? stamps + (size_t)blockIdx.x * kStampsPerGroup : nullptr;
  if (my_stamps && tid == 0) my_stamps[kStampStart] = wall_clock64();
  const uint32_t epoch = ld_relaxed(&ctrl[kCtrlMoveEpoch]);
  uint64_t* __restrict__ st = mstatus + (size_t)(epoch & 1u) * kMaxSlices;
  {
    uint64_t* other = mstatus + (size_t)((epoch & 1u) ^ 1u) * kMaxSlices;
    for (uint32_t w = slice * WG + tid; w < (uint32_t)kMaxSlices; w += ns * WG) other[w] = 0ull;
  }
  move_pair_core<T, WG, IPT, COH, COH>(A, R, k, sZ, L, ns, bits, st, slots, slot_cap, ctrl, stats, early, slice,
                                       backoff, epoch, my_stamps);
  // the last workgroup to leave (all of them have read the epoch by then)
  // starts the next call's epoch.  With err_out, it also hands the call's
  // error bits to the caller's word: every bit a workgroup sets is set by its
  // thread 0, whose returning atomics have completed before its exit count.
  if (tid == 0) {
    if (err_out) wait_all_memory();
    if (add_relaxed(&ctrl[kCtrlMoveExit], 1u) == ns - 1u) {
      st_relaxed(&ctrl[kCtrlMoveExit], 0u);
      st_relaxed(&ctrl[kCtrlMoveEpoch], epoch + 1u);
      if (err_out) {
        const uint32_t e = ld_relaxed(&ctrl[kCtrlError]);
        if (e) or_relaxed(err_out, e);
      }
    }
    if (my_stamps) my_stamps[kStampExit] = wall_clock64();
  }
}

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_move_pair(T* __restrict__ A, const int32_t* __restrict__ R, uint32_t k,
                                                    uint32_t sZ, uint32_t L, uint32_t ns,
                                                    uint32_t* __restrict__ bits, uint64_t* __restrict__ mstatus,
                                                    uint64_t* __restrict__ slots, uint32_t slot_cap,
                                                    uint32_t* __restrict__ ctrl, int stats, int early,
                                                    uint32_t* __restrict__ err_out, uint64_t* __restrict__ stamps) {
  move_pair_body<T, WG, IPT, false>(A, R, k, sZ, L, ns, bits, mstatus, slots, slot_cap, ctrl, stats, early,
                                    blockIdx.x, 0u, err_out, stamps);
}

// Variant 6: the mark step in the same launch.  Workgroups [0, nm) each mark
// a tile of `tile` removals and count themselves done, and the last of them
// raises a ready word per XCD; the move-and-pair slices come after them in
// dispatch order (so every mark workgroup is running or done when a slice
// waits), prefetch their first step and wait for their XCD's ready word.
constexpr int kFusedMarkTileLog = 12;  // default tile: 4096 removals
constexpr uint32_t kMarkWaitBackoff = 16;  // longest pause between polls of the ready word: 16 x 64 cycles

// One mark workgroup of the one-launch form (id < nm): mark its tile of the
// removals, then count itself done.
template <int WG>
__device__ __forceinline__ void mark_group_body(const int32_t* __restrict__ R, uint32_t k, uint32_t sZ,
                                                uint32_t* __restrict__ bits, uint32_t* __restrict__ ctrl, int stats,
                                                uint32_t nm, uint32_t tile, int sorted,
                                                uint64_t* __restrict__ stamps) {
  uint64_t* const my_stamps = stamps ? stamps + (size_t)blockIdx.x * kStampsPerGroup : nullptr;
  if (my_stamps && threadIdx.x == 0) my_stamps[kStampStart] = wall_clock64();
  const uint32_t lo = blockIdx.x * tile, hi = min(lo + tile, k);
  if (!sorted || (uint32_t)R[hi - 1] >= sZ) {  // a sorted chunk below the red zone has no marks
    // a whole tile's loads in flight at once (16 per thread for 4,096 removals and 256 threads)
    if (stats)
      mark_body<true, 16>(R + lo, hi - lo, k, sZ, bits, (uint64_t*)&ctrl[kCtrlStatTail], threadIdx.x, WG);
    else
      mark_body<false, 16>(R + lo, hi - lo, k, sZ, bits, nullptr, threadIdx.x, WG);
  }
  // Done counts in two levels: the mark workgroups with ids equal mod 8
  // share a group counter, the last of each group adds one to the global
  // count, and the last group resets it and raises the 8 ready words (one
  // per XCD, so that the waiting slices poll 8 lines).  One counter for all
  // of them serialises: same-address atomics take about 11 ns each.
  wait_all_memory();
  __syncthreads();
  if (my_stamps && threadIdx.x == 0) my_stamps[kStampMarked] = wall_clock64();
  if (threadIdx.x == 0) {
    const uint32_t x = blockIdx.x & 7u;
    const uint32_t gsize = (nm - x + 7u) / 8u;  // mark workgroups with id = x mod 8
    if (add_relaxed(&ctrl[kCtrlMarkGroup + 32 * x], 1u) == gsize - 1u) {
      st_relaxed(&ctrl[kCtrlMarkGroup + 32 * x], 0u);
      const uint32_t groups = nm < 8u ? nm : 8u;
      if (add_relaxed(&ctrl[kCtrlMarkDone], 1u) == groups - 1u) {
        st_relaxed(&ctrl[kCtrlMarkDone], 0u);
        const uint32_t ready = ld_relaxed(&ctrl[kCtrlMoveEpoch]) + 1u;  // no slice can have finished yet
        for (int r = 0; r < 8; ++r) st_relaxed(&ctrl[kCtrlMarkReady + 32 * r], ready);
        if (my_stamps) my_stamps[kStampRaised] = wall_clock64();
      }
    }
  }
}

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_mark_move_pair(T* __restrict__ A, const int32_t* __restrict__ R, uint32_t k,
                                                         uint32_t sZ, uint32_t L, uint32_t ns,
                                                         uint32_t* __restrict__ bits,
                                                         uint64_t* __restrict__ mstatus,
                                                         uint64_t* __restrict__ slots, uint32_t slot_cap,
                                                         uint32_t* __restrict__ ctrl, int stats, uint32_t nm,
                                                         uint32_t tile, uint32_t backoff, int sorted, int early,
                                                         uint32_t* __restrict__ err_out,
                                                         uint64_t* __restrict__ stamps) {
  if (blockIdx.x < nm) {
    mark_group_body<WG>(R, k, sZ, bits, ctrl, stats, nm, tile, sorted, stamps);
    return;
  }
  move_pair_body<T, WG, IPT, true>(A, R, k, sZ, L, ns, bits, mstatus, slots, slot_cap, ctrl, stats, early,
                                   blockIdx.x - nm, backoff, err_out, stamps);
}

// The one-launch form with n and k read from device memory (rz_remove_sized;
// the reference's TestGPUWide dereferences sizeA and sizeR in the kernel,
// redzone.hxx:457-465).  The grid is sized on the host for k_cap removals:
// nm_cap mark workgroups, then ns_cap = min(cap, ceil(k_cap / step)) slices.
// Each workgroup derives this call's shape from the device k with the host's
// slice rule (L = whole steps so that at most `cap` slices cover k), so
// L(k) <= L(k_cap) fits the launch's LDS and ns(k) <= ns_cap; the workgroups
// beyond nm(k) marks and ns(k) slices leave at once and take part in no count.
// Sizes outside 0 <= k <= n <= n_cap, k <= k_cap move nothing and set
// kErrSizes.
template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_mark_move_pair_sized(T* __restrict__ A, const int32_t* __restrict__ R,
                                                               const int32_t* __restrict__ d_n,
                                                               const int32_t* __restrict__ d_k, uint32_t n_cap,
                                                               uint32_t k_cap, uint32_t cap, uint32_t nm_cap,
                                                               uint32_t* __restrict__ bits,
                                                               uint64_t* __restrict__ mstatus,
                                                               uint64_t* __restrict__ slots,
                                                               uint32_t* __restrict__ ctrl, int stats, uint32_t tile,
                                                               uint32_t backoff, int sorted, int early) {
  const int64_t n = d_n[0], k = d_k[0];
  if (k < 0 || k > n || n > (int64_t)n_cap || k > (int64_t)k_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) or_relaxed(&ctrl[kCtrlError], kErrSizes);
    return;
  }
  if (k == 0 || k == n) return;  // nothing moves
  const uint32_t uk = (uint32_t)k, sZ = (uint32_t)(n - k);
  constexpr uint64_t step = (uint64_t)WG * IPT;
  const uint64_t span = step * cap;
  const uint32_t L = (uint32_t)(((uk + span - 1) / span) * step);
  const uint32_t ns = (uk + L - 1) / L;
  const uint32_t nm = (uint32_t)(((uint64_t)uk + tile - 1) / tile);
  if (blockIdx.x < nm_cap) {
    if (blockIdx.x < nm) mark_group_body<WG>(R, uk, sZ, bits, ctrl, stats, nm, tile, sorted, nullptr);
    return;
  }
  const uint32_t slice = blockIdx.x - nm_cap;
  if (slice >= ns) return;
  move_pair_body<T, WG, IPT, true>(A, R, uk, sZ, L, ns, bits, mstatus, slots, uk / 2 + 1, ctrl, stats, early, slice,
                                   backoff, nullptr, nullptr);
}

// ---------------------------------------------------------------------------
// cooperative single launch: the three steps separated by in-kernel grid
// barriers built from relaxed atomics (no agent-scope fences; everything a
// later step reads across workgroups was written coherently)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void grid_barrier(uint32_t* __restrict__ ctrl, uint64_t target) {
  wait_all_memory();  // this thread's coherent stores have been acknowledged
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t* cnt = (uint64_t*)&ctrl[kCtrlBarCount];
    add_relaxed(cnt, 1ull);
    uint32_t spins = 0;
    const uint32_t cap = spin_cap(ctrl, 1u << 22);
    while (ld_relaxed(cnt) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > cap) {
        or_relaxed(&ctrl[kCtrlError], kErrBarrierTimeout);
        break;
      }
    }
  }
  __syncthreads();
}

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_coop(T* __restrict__ A, const int32_t* __restrict__ R,
                                               int32_t* __restrict__ Rd, uint32_t k, uint32_t sZ, uint32_t L,
                                               uint32_t ns, uint32_t* __restrict__ bits,
                                               uint64_t* __restrict__ counts, uint32_t* __restrict__ ctrl,
                                               int stats) {
  __shared__ MoveShared<WG, IPT> msh;
  __shared__ PairShared<WG, kMaxSlices> psh;
  const uint64_t base = ld_relaxed((const uint64_t*)&ctrl[kCtrlBarEpoch]);
  const uint32_t G = gridDim.x;
  if (stats)
    mark_body<true>(R, k, k, sZ, bits, (uint64_t*)&ctrl[kCtrlStatTail], blockIdx.x * WG + threadIdx.x, G * WG);
  else
    mark_body<false>(R, k, k, sZ, bits, nullptr, blockIdx.x * WG + threadIdx.x, G * WG);
  grid_barrier(ctrl, base + G);
  // every workgroup has read `base` before arriving: publish the next launch's base
  if (blockIdx.x == 0 && threadIdx.x == 0) st_relaxed((uint64_t*)&ctrl[kCtrlBarEpoch], base + 2ull * G);
  for (uint32_t s = blockIdx.x; s < ns; s += G) move_slice<T, WG, IPT, true>(A, R, Rd, k, sZ, L, s, bits, counts, msh);
  grid_barrier(ctrl, base + 2ull * G);
  pair_body<T, WG, kMaxSlices, true>(A, Rd, k, sZ, L, ns, counts, ctrl, stats, blockIdx.x, G, psh);
}

// ---------------------------------------------------------------------------
// ordered single launch: workgroups take their role from their id,
// [0, nm) mark tiles of LM removals, [nm, nm+ns) move slices, the rest pair
// shares.  Workgroups are dispatched in id order and each waits only on lower
// ids, so the launch needs neither a cooperative launch nor co-residency.
// Every workgroup does one item and exits, so the random stores of one slice
// never hold up the loads of the next (a persistent workgroup's vmcnt would).
// The last workgroup to leave resets the three counters.
// ---------------------------------------------------------------------------
enum OrderedWord : int { kOrdMark = 0, kOrdMove = 32, kOrdExit = 64 };
constexpr uint32_t kOrdMarkTile = 4096;
constexpr int kOrdCtlOffset = 128;  // words into the context's queue block

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_ordered(T* __restrict__ A, const int32_t* __restrict__ R,
                                                  int32_t* __restrict__ Rd, uint32_t k, uint32_t sZ, uint32_t L,
                                                  uint32_t ns, uint32_t nm, uint32_t* __restrict__ bits,
                                                  uint64_t* __restrict__ counts, uint32_t* __restrict__ ctrl,
                                                  uint32_t* __restrict__ q, int stats) {
  __shared__ MoveShared<WG, IPT> msh;
  __shared__ uint32_t s_last;
  extern __shared__ uint64_t dyn[];  // pairing prefix: ns + 1 entries, then WG/64 + 1 scan cells
  const uint32_t b = blockIdx.x, tid = threadIdx.x;
  if (b < nm) {
    const uint32_t lo = b * kOrdMarkTile, hi = min(lo + kOrdMarkTile, k);
    if (stats)
      mark_body<true>(R + lo, hi - lo, k, sZ, bits, (uint64_t*)&ctrl[kCtrlStatTail], tid, WG);
    else
      mark_body<false>(R + lo, hi - lo, k, sZ, bits, nullptr, tid, WG);
    signal_done(&q[kOrdMark]);
  } else if (b < nm + ns) {
    wait_count(&q[kOrdMark], nm, ctrl);
    move_slice<T, WG, IPT, true>(A, R, Rd, k, sZ, L, b - nm, bits, counts, msh, &q[kOrdMove]);
  } else {
    wait_count(&q[kOrdMove], ns, ctrl);
    pair_body_lds<T, WG, true>(A, Rd, k, sZ, L, ns, counts, ctrl, stats, b - nm - ns, gridDim.x - nm - ns, dyn,
                               dyn + ns + 1);
  }
  if (tid == 0) {
    wait_all_memory();
    s_last = add_relaxed(&q[kOrdExit], 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (s_last && tid == 0) {
    st_relaxed(&q[kOrdMark], 0u);
    st_relaxed(&q[kOrdMove], 0u);
    st_relaxed(&q[kOrdExit], 0u);
  }
}

// ---------------------------------------------------------------------------
// launch
// ---------------------------------------------------------------------------
template <typename T, int WG, int IPT>
static rz_status ordered_launch(rz_ctx* ctx, T* A, int32_t* R, int32_t* Rd, uint32_t k, uint32_t sZ, uint32_t L,
                                uint32_t ns, uint32_t* bits, uint64_t* counts, uint32_t* ctrl, int stats,
                                hipStream_t s) {
  const uint32_t nm = (k + kOrdMarkTile - 1) / kOrdMarkTile;
  uint32_t np = ns / 8;
  if (np < 1) np = 1;
  if (np > (uint32_t)ctx->cu_count) np = (uint32_t)ctx->cu_count;
  const size_t lds = ((size_t)ns + 1 + WG / kWave + 1) * sizeof(uint64_t);
  rz_status st;
  if ((st = ensure(ctx, ctx->queue, 4096, s)) != RZ_OK) return st;
  uint32_t* q = (uint32_t*)ctx->queue.ptr + kOrdCtlOffset;
  hipLaunchKernelGGL((rz_ordered<T, WG, IPT>), dim3(nm + ns + np), dim3(WG), lds, s, A, R, Rd, k, sZ, L, ns, nm,
                     bits, counts, ctrl, q, stats);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

template <typename T>
static rz_status dispatch_ordered(int wg, int ipt, rz_ctx* ctx, T* A, int32_t* R, int32_t* Rd, uint32_t k,
                                  uint32_t sZ, uint32_t L, uint32_t ns, uint32_t* bits, uint64_t* counts,
                                  uint32_t* ctrl, int stats, hipStream_t s) {
#define RZ_ORD_CASE(W, I) \
  if (wg == W && ipt == I) return ordered_launch<T, W, I>(ctx, A, R, Rd, k, sZ, L, ns, bits, counts, ctrl, stats, s);
  RZ_ORD_CASE(256, 2) RZ_ORD_CASE(256, 4) RZ_ORD_CASE(256, 8) RZ_ORD_CASE(256, 16)
  RZ_ORD_CASE(512, 2) RZ_ORD_CASE(512, 4) RZ_ORD_CASE(512, 8) RZ_ORD_CASE(512, 16)
#undef RZ_ORD_CASE
  return RZ_EINVAL;
}
template <typename T, int WG, int IPT>
static void launch_move(dim3 grid, hipStream_t s, T* A, const int32_t* R, int32_t* Rd, uint32_t k, uint32_t sZ,
                        uint32_t L, uint32_t* bits, uint64_t* counts) {
  hipLaunchKernelGGL((rz_move<T, WG, IPT>), grid, dim3(WG), 0, s, A, R, Rd, k, sZ, L, bits, counts);
}

template <typename T>
static bool dispatch_move(int wg, int ipt, dim3 grid, hipStream_t s, T* A, const int32_t* R, int32_t* Rd,
                          uint32_t k, uint32_t sZ, uint32_t L, uint32_t* bits, uint64_t* counts) {
#define RZ_MOVE_CASE(W, I) \
  if (wg == W && ipt == I) { launch_move<T, W, I>(grid, s, A, R, Rd, k, sZ, L, bits, counts); return true; }
  RZ_MOVE_CASE(256, 2) RZ_MOVE_CASE(256, 4) RZ_MOVE_CASE(256, 8) RZ_MOVE_CASE(256, 16)
  RZ_MOVE_CASE(512, 2) RZ_MOVE_CASE(512, 4) RZ_MOVE_CASE(512, 8) RZ_MOVE_CASE(512, 16)
#undef RZ_MOVE_CASE
  return false;
}

template <typename T>
static bool dispatch_move_pair(int wg, int ipt, hipStream_t s, T* A, const int32_t* R, uint32_t k, uint32_t sZ,
                               uint32_t L, uint32_t ns, uint32_t* bits, uint64_t* mstatus, uint64_t* slots,
                               uint32_t slot_cap, uint32_t* ctrl, int stats, int early, uint32_t* err_out = nullptr,
                               uint64_t* stamps = nullptr) {
  const size_t lds = (size_t)L * 4;
#define RZ_MP_CASE(W, I)                                                                                     \
  if (wg == W && ipt == I) {                                                                                 \
    hipLaunchKernelGGL((rz_move_pair<T, W, I>), dim3(ns), dim3(W), lds, s, A, R, k, sZ, L, ns, bits, mstatus, \
                       slots, slot_cap, ctrl, stats, early, err_out, stamps);                                \
    return true;                                                                                             \
  }
  RZ_MP_CASE(256, 2) RZ_MP_CASE(256, 4) RZ_MP_CASE(256, 8) RZ_MP_CASE(256, 16)
  RZ_MP_CASE(512, 2) RZ_MP_CASE(512, 4) RZ_MP_CASE(512, 8) RZ_MP_CASE(512, 16)
#undef RZ_MP_CASE
  return false;
}

static int clamp_tile_log(int tile_log) {
  return tile_log == 0 ? kFusedMarkTileLog : (tile_log < 8 ? 8 : (tile_log > 16 ? 16 : tile_log));
}
// mark workgroups of the one-launch form for k removals
static uint32_t mark_groups(int64_t k, int tile_log) {
  const uint32_t tile = 1u << clamp_tile_log(tile_log);
  return (uint32_t)((k + tile - 1) / tile);
}

// Variant 6: nm mark workgroups, then the ns move-and-pair slices, one grid.
template <typename T>
static bool dispatch_mark_move_pair(int wg, int ipt, hipStream_t s, T* A, const int32_t* R, uint32_t k, uint32_t sZ,
                                    uint32_t L, uint32_t ns, uint32_t* bits, uint64_t* mstatus, uint64_t* slots,
                                    uint32_t slot_cap, uint32_t* ctrl, int stats, int tile_log, int sorted,
                                    int early, uint32_t* err_out = nullptr, uint64_t* stamps = nullptr) {
  const uint32_t bo = kMarkWaitBackoff;
  const size_t lds = (size_t)L * 4;
  tile_log = clamp_tile_log(tile_log);
  const uint32_t tile = 1u << tile_log;
  const uint32_t nm = (k + tile - 1) / tile;
#define RZ_MMP_CASE(W, I)                                                                                      \
  if (wg == W && ipt == I) {                                                                                   \
    hipLaunchKernelGGL((rz_mark_move_pair<T, W, I>), dim3(nm + ns), dim3(W), lds, s, A, R, k, sZ, L, ns, bits, \
                       mstatus, slots, slot_cap, ctrl, stats, nm, tile, bo, sorted, early, err_out, stamps);   \
    return true;                                                                                               \
  }
  RZ_MMP_CASE(256, 2) RZ_MMP_CASE(256, 4) RZ_MMP_CASE(256, 8) RZ_MMP_CASE(256, 16)
  RZ_MMP_CASE(512, 2) RZ_MMP_CASE(512, 4) RZ_MMP_CASE(512, 8) RZ_MMP_CASE(512, 16)
#undef RZ_MMP_CASE
  return false;
}

// The fused form keeps a slice's orphan list in LDS (4 bytes per pair); a
// slice too long for the device's per-workgroup LDS runs the three-kernel form.
static bool fused_fits(const rz_ctx* ctx, const rz_tile_config& cfg, uint32_t L) {
  const size_t fixed = 4096;  // static LDS of the kernel, rounded up
  return (size_t)L * 4 + fixed <= (size_t)ctx->lds_per_block && cfg.wg_size * cfg.items <= (int)L;
}

void slice_shape(int64_t k, const rz_tile_config& cfg, int max_slices_cap, uint32_t* L, uint32_t* ns) {
  const int64_t step = (int64_t)cfg.wg_size * cfg.items;
  int64_t cap = cfg.max_slices < 1 ? 1 : cfg.max_slices;
  if (cap > max_slices_cap) cap = max_slices_cap;
  const int64_t steps = (k + step * cap - 1) / (step * cap);
  *L = (uint32_t)(steps * step);
  *ns = (uint32_t)((k + *L - 1) / *L);
}

template <typename T, int WG, int IPT>
static rz_status coop_launch(rz_ctx* ctx, T* A, int32_t* R, int32_t* Rd, uint32_t k, uint32_t sZ, uint32_t L,
                             uint32_t ns, uint32_t* bits, uint64_t* counts, uint32_t* ctrl, int stats,
                             hipStream_t s) {
  int per_cu = 0;
  RZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rz_coop<T, WG, IPT>, WG, 0));
  if (per_cu < 1) return RZ_ECOOP;
  const uint32_t resident = (uint32_t)per_cu * (uint32_t)ctx->cu_count;
  const uint32_t G = ns < resident ? ns : resident;
  void* args[] = {&A, &R, &Rd, &k, &sZ, &L, &ns, &bits, &counts, &ctrl, &stats};
  const hipError_t e = hipLaunchCooperativeKernel((const void*)rz_coop<T, WG, IPT>, dim3(G), dim3(WG), args, 0, s);
  if (e == hipErrorCooperativeLaunchTooLarge) return RZ_ECOOP;
  if (e != hipSuccess) return RZ_EHIP;
  return RZ_OK;
}

template <typename T>
static rz_status dispatch_coop(int wg, int ipt, rz_ctx* ctx, T* A, int32_t* R, int32_t* Rd, uint32_t k,
                               uint32_t sZ, uint32_t L, uint32_t ns, uint32_t* bits, uint64_t* counts,
                               uint32_t* ctrl, int stats, hipStream_t s) {
#define RZ_COOP_CASE(W, I) \
  if (wg == W && ipt == I) return coop_launch<T, W, I>(ctx, A, R, Rd, k, sZ, L, ns, bits, counts, ctrl, stats, s);
  RZ_COOP_CASE(256, 2) RZ_COOP_CASE(256, 4) RZ_COOP_CASE(256, 8) RZ_COOP_CASE(256, 16)
  RZ_COOP_CASE(512, 2) RZ_COOP_CASE(512, 4) RZ_COOP_CASE(512, 8) RZ_COOP_CASE(512, 16)
#undef RZ_COOP_CASE
  return RZ_EINVAL;
}

template <typename T>
static rz_status plain_impl(rz_ctx* ctx, T* A, int64_t n, int32_t* R, int64_t k, uint32_t flags,
                            const rz_tile_config& cfg, bool coop, hipStream_t s, uint32_t* ext_bits) {
  const uint32_t sZ = (uint32_t)(n - k);
  uint32_t L, ns;
  slice_shape(k, cfg, kMaxSlices, &L, &ns);
  rz_status st;
  ctx->stamp_groups = 0;  // stage stamps only from a fused call with RZ_PROFILE
  if (!ext_bits && (st = ensure(ctx, ctx->tail_bits, (size_t)((k + 63) / 64) * 8, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->counts, (size_t)ns * 8, s)) != RZ_OK) return st;
  int32_t* Rd = R;
  if (flags & RZ_PRESERVE_REMOVALS) {
    if ((st = ensure(ctx, ctx->orphans, (size_t)k * 4, s)) != RZ_OK) return st;
    Rd = (int32_t*)ctx->orphans.ptr;
  }
  uint32_t* bits = ext_bits ? ext_bits : (uint32_t*)ctx->tail_bits.ptr;
  uint64_t* counts = (uint64_t*)ctx->counts.ptr;
  uint32_t* ctrl = (uint32_t*)ctx->ctrl.ptr;
  const int stats = (flags & RZ_STATS) ? 1 : 0;
  if (stats) {  // RZ_STATS reports the counts of this call only (the error word is cleared when it is read)
    if ((st = fill_u32(ctrl + kCtrlStatTail, 0u, 2, s)) != RZ_OK) return st;
    if ((st = fill_u32(ctrl + kCtrlStatOrphans, 0u, 2, s)) != RZ_OK) return st;
  }
  if ((st = prof_mark(ctx, flags, 0, s)) != RZ_OK) return st;
  if (coop || cfg.variant == 3) {
    if ((st = prof_mark(ctx, flags, 1, s)) != RZ_OK) return st;
    if (coop)
      st = dispatch_coop<T>(cfg.wg_size, cfg.items, ctx, A, R, Rd, (uint32_t)k, sZ, L, ns, bits, counts, ctrl, stats,
                            s);
    else
      st = dispatch_ordered<T>(cfg.wg_size, cfg.items, ctx, A, R, Rd, (uint32_t)k, sZ, L, ns, bits, counts, ctrl,
                               stats, s);
    if (st != RZ_OK) return st;
    if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
    return prof_mark(ctx, flags, 3, s);
  }
  const bool fused = (cfg.variant == 4 || cfg.variant == 6) && fused_fits(ctx, cfg, L);
  const bool one_launch = fused && cfg.variant == 6;  // marks inside the fused kernel
  if (fused) {
    if ((st = ensure(ctx, ctx->mstatus, 2 * (size_t)kMaxSlices * 8, s)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->slots, ((size_t)k / 2 + 1) * 8, s)) != RZ_OK) return st;
  }
  const bool premarked = (flags & kFlagPremarked) != 0;
  const bool dense_marks = !premarked && dense_bucket_marks(cfg, k, flags);
  const int64_t g1w = (k + 256 * 4 - 1) / (256 * 4);
  const dim3 g1((uint32_t)(g1w < (int64_t)ctx->cu_count * 8 ? g1w : ctx->cu_count * 8));
  if (premarked || dense_marks || one_launch) {
    // the caller's producer set the tail marks, or the launch below sets them itself
  } else if (stats)
    hipLaunchKernelGGL(rz_mark_tail<true>, g1, dim3(256), 0, s, R, (uint32_t)k, sZ, bits,
                       (uint64_t*)(ctrl + kCtrlStatTail));
  else if (flags & RZ_SORTED)
    hipLaunchKernelGGL(rz_mark_tail_sorted, dim3((uint32_t)((k + kSortedChunk - 1) / kSortedChunk)), dim3(256), 0,
                       s, R, (uint32_t)k, sZ, bits);
  else
    hipLaunchKernelGGL(rz_mark_tail<false>, g1, dim3(256), 0, s, R, (uint32_t)k, sZ, bits, (uint64_t*)nullptr);
  RZ_HIP(hipGetLastError());
  if ((st = prof_mark(ctx, flags, 1, s)) != RZ_OK) return st;
  if (cfg.variant == 5) {
    if ((st = dense_moves<T>(ctx, A, n, R, k, cfg, bits, ctrl, stats, dense_marks, s)) != RZ_OK) {
      // the marks are in but no move step will clear them: leave the bitmap zero
      if (!premarked) (void)fill_u32(bits, 0u, (size_t)((k + 31) / 32), s);
      return st;
    }
    if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
    return prof_mark(ctx, flags, 3, s);
  }
  uint64_t* stamps = nullptr;
  if (fused && (flags & RZ_PROFILE)) {
    const uint32_t groups = ns + (one_launch ? mark_groups(k, cfg.slice_steps) : 0u);
    if ((st = ensure(ctx, ctx->stamps, (size_t)groups * kStampsPerGroup * 8, s)) != RZ_OK) return st;
    stamps = (uint64_t*)ctx->stamps.ptr;
    if ((st = fill_u32((uint32_t*)stamps, 0u, (size_t)groups * kStampsPerGroup * 2, s)) != RZ_OK) return st;
    ctx->stamp_groups = groups;
    ctx->stamp_marks = one_launch ? mark_groups(k, cfg.slice_steps) : 0u;
  }
  if (one_launch) {
    if (!dispatch_mark_move_pair<T>(cfg.wg_size, cfg.items, s, A, R, (uint32_t)k, sZ, L, ns, bits,
                                    (uint64_t*)ctx->mstatus.ptr, (uint64_t*)ctx->slots.ptr, (uint32_t)(k / 2 + 1),
                                    ctrl, stats, cfg.slice_steps, (flags & RZ_SORTED) ? 1 : 0,
                                    (flags & RZ_ANY_PAIRING) ? 1 : 0, nullptr, stamps))
      return RZ_EINVAL;
    RZ_HIP(hipGetLastError());
    if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
    return prof_mark(ctx, flags, 3, s);
  }
  if (fused) {
    if (!dispatch_move_pair<T>(cfg.wg_size, cfg.items, s, A, R, (uint32_t)k, sZ, L, ns, bits,
                               (uint64_t*)ctx->mstatus.ptr, (uint64_t*)ctx->slots.ptr, (uint32_t)(k / 2 + 1), ctrl,
                               stats, (flags & RZ_ANY_PAIRING) ? 1 : 0, nullptr, stamps))
      return RZ_EINVAL;
    RZ_HIP(hipGetLastError());
    if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
    return prof_mark(ctx, flags, 3, s);
  }
  if (!dispatch_move<T>(cfg.wg_size, cfg.items, dim3(ns), s, A, R, Rd, (uint32_t)k, sZ, L, bits, counts))
    return RZ_EINVAL;
  RZ_HIP(hipGetLastError());
  if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
  const dim3 g3((uint32_t)ctx->cu_count);
  hipLaunchKernelGGL(rz_pair<T>, g3, dim3(256), 0, s, A, Rd, (uint32_t)k, sZ, L, ns, counts, ctrl, stats);
  RZ_HIP(hipGetLastError());
  return prof_mark(ctx, flags, 3, s);
}

// ---------------------------------------------------------------------------
// the fused forms on scratch the caller provides (rz_remove_noalloc)
// ---------------------------------------------------------------------------
// Clears what a fused call needs at zero: the tail bitmap words of this k, the
// error word, the exit counter and the mark-done counters.  A kernel, not a memset, so that the
// whole call can be captured in a graph.
__global__ __launch_bounds__(256) void rz_static_prologue(uint32_t* __restrict__ bits, uint32_t nwords,
                                                          uint32_t* __restrict__ ctrl) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  for (uint32_t w = i; w < nwords; w += gridDim.x * 256u) bits[w] = 0u;
  if (i == 0) {
    ctrl[kCtrlError] = 0u;
    ctrl[kCtrlMoveExit] = 0u;
    ctrl[kCtrlMarkDone] = 0u;
  }
  if (i < 8) ctrl[kCtrlMarkGroup + 32 * i] = 0u;
}

template <typename T>
static rz_status fused_static_impl(T* A, int64_t n, const int32_t* R, int64_t k, uint32_t flags,
                                   const rz_tile_config& cfg, const FusedScratch& fs, hipStream_t s) {
  const uint32_t sZ = (uint32_t)(n - k);
  uint32_t L, ns;
  slice_shape(k, cfg, kMaxSlices, &L, &ns);
  if ((size_t)L * 4 + 4096 > 65536 || cfg.wg_size * cfg.items > (int)L) return RZ_EINVAL;
  const uint32_t nwords = (uint32_t)((k + 31) / 32);
  const uint32_t g0 = (nwords + 255) / 256;
  hipLaunchKernelGGL(rz_static_prologue, dim3(g0 < 1024 ? (g0 ? g0 : 1) : 1024), dim3(256), 0, s, fs.bits, nwords,
                     fs.ctrl);
  RZ_HIP(hipGetLastError());
  const int early = (flags & RZ_ANY_PAIRING) ? 1 : 0;
  const uint32_t slot_cap = (uint32_t)(k / 2 + 1);
  if (cfg.variant == 6) {
    if (!dispatch_mark_move_pair<T>(cfg.wg_size, cfg.items, s, A, R, (uint32_t)k, sZ, L, ns, fs.bits, fs.mstatus,
                                    fs.slots, slot_cap, fs.ctrl, 0, cfg.slice_steps, (flags & RZ_SORTED) ? 1 : 0,
                                    early, fs.err_out))
      return RZ_EINVAL;
  } else {
    if (flags & RZ_SORTED)
      hipLaunchKernelGGL(rz_mark_tail_sorted, dim3((uint32_t)((k + kSortedChunk - 1) / kSortedChunk)), dim3(256), 0,
                         s, R, (uint32_t)k, sZ, fs.bits);
    else {
      const int64_t g1w = (k + 256 * 4 - 1) / (256 * 4);
      hipLaunchKernelGGL(rz_mark_tail<false>, dim3((uint32_t)(g1w < 2048 ? g1w : 2048)), dim3(256), 0, s, R,
                         (uint32_t)k, sZ, fs.bits, (uint64_t*)nullptr);
    }
    RZ_HIP(hipGetLastError());
    if (!dispatch_move_pair<T>(cfg.wg_size, cfg.items, s, A, R, (uint32_t)k, sZ, L, ns, fs.bits, fs.mstatus,
                               fs.slots, slot_cap, fs.ctrl, 0, early, fs.err_out))
      return RZ_EINVAL;
  }
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

rz_status launch_fused_static(void* data, int64_t n, uint32_t esz, const int32_t* R, int64_t k, uint32_t flags,
                              const rz_tile_config& cfg, const FusedScratch& fs, hipStream_t s) {
  if (cfg.variant != 4 && cfg.variant != 6) return RZ_EINVAL;
  switch (esz) {
    case 4: return fused_static_impl<uint32_t>((uint32_t*)data, n, R, k, flags, cfg, fs, s);
    case 8: return fused_static_impl<uint64_t>((uint64_t*)data, n, R, k, flags, cfg, fs, s);
    case 16: return fused_static_impl<uint4>((uint4*)data, n, R, k, flags, cfg, fs, s);
    default: return RZ_EINVAL;
  }
}

// Whether a launch shape can run, checked before any kernel of a call is
// launched: a call that marked the tail and then refused its move step would
// leave the bitmap set for the next call.
static bool plain_shape_ok(const rz_ctx* ctx, const rz_tile_config& cfg, uint32_t esz, int64_t n, int64_t k,
                           uint32_t flags) {
  const bool wg_ok = cfg.wg_size == 256 || cfg.wg_size == 512;
  switch (cfg.variant) {
    case 0: case 1: case 3: case 4: case 6:
      return wg_ok && (cfg.items == 2 || cfg.items == 4 || cfg.items == 8 || cfg.items == 16);
    case 2: return wg_ok && (cfg.items == 4 || cfg.items == 8);
    case 5: return dense_shape_ok(ctx, cfg, esz, n, k, flags);
    default: return false;
  }
}

rz_status launch_plain(rz_ctx* ctx, void* data, int64_t n, uint32_t esz, int32_t* R, int64_t k, uint32_t flags,
                       const rz_tile_config& cfg_in, hipStream_t s, uint32_t* ext_bits) {
  rz_tile_config cfg = cfg_in;
  const bool premarked = (flags & kFlagPremarked) != 0;
  // the single-launch forms include their own mark step: with the marks set
  // by the caller, run their shape as the two-kernel fused form
  if (premarked && ((cfg.variant >= 1 && cfg.variant <= 3) || cfg.variant == 6)) cfg.variant = 4;
  if (!plain_shape_ok(ctx, cfg, esz, n, k, flags)) return RZ_EINVAL;
  if (cfg.variant == 2) return launch_queue_single(ctx, data, n, esz, R, k, flags, cfg, s);
  const bool coop = !premarked && (cfg.variant == 1 || (flags & RZ_COOPERATIVE));
  switch (esz) {
    case 4: return plain_impl<uint32_t>(ctx, (uint32_t*)data, n, R, k, flags, cfg, coop, s, ext_bits);
    case 8: return plain_impl<uint64_t>(ctx, (uint64_t*)data, n, R, k, flags, cfg, coop, s, ext_bits);
    case 16: return plain_impl<uint4>(ctx, (uint4*)data, n, R, k, flags, cfg, coop, s, ext_bits);
    default: return RZ_EINVAL;
  }
}

// ---------------------------------------------------------------------------
// sizes in device memory (rz_remove_sized)
// ---------------------------------------------------------------------------
// The one-launch shape for up to k_cap removals: the context's tuned shape for
// k_cap when it is a fused form whose slices fit LDS at k_cap, else 256 x 8 with
// up to kMaxSlices slices.  false when no fused shape fits k_cap.
bool sized_shape(const rz_ctx* ctx, uint32_t esz, int64_t k_cap, rz_tile_config* out) {
  rz_tile_config c = ctx->table[esz_index(esz)][log2_floor(k_cap)];
  uint32_t L, ns;
  const bool tuned = (c.variant == 4 || c.variant == 6) && (c.wg_size == 256 || c.wg_size == 512) &&
                     (c.items == 2 || c.items == 4 || c.items == 8 || c.items == 16);
  if (tuned) {
    if (c.variant == 4) c.slice_steps = 0;  // the default mark tile
    c.variant = 6;
    slice_shape(k_cap, c, kMaxSlices, &L, &ns);
    if (fused_fits(ctx, c, L)) {
      *out = c;
      return true;
    }
  }
  c = rz_tile_config{256, 8, kMaxSlices, 6, 0, 0};
  slice_shape(k_cap, c, kMaxSlices, &L, &ns);
  if (!fused_fits(ctx, c, L)) return false;
  *out = c;
  return true;
}

template <typename T>
static rz_status sized_impl(rz_ctx* ctx, T* A, int64_t n_cap, const int32_t* R, int64_t k_cap, const int32_t* d_n,
                            const int32_t* d_k, uint32_t flags, const rz_tile_config& cfg, hipStream_t s) {
  uint32_t L_cap, ns_unused;
  slice_shape(k_cap, cfg, kMaxSlices, &L_cap, &ns_unused);
  const int64_t step = (int64_t)cfg.wg_size * cfg.items;
  const uint32_t cap = (uint32_t)(cfg.max_slices < 1 ? 1 : (cfg.max_slices > kMaxSlices ? kMaxSlices : cfg.max_slices));
  const int64_t ns_k = (k_cap + step - 1) / step;
  const uint32_t ns_cap = (uint32_t)(ns_k < (int64_t)cap ? ns_k : (int64_t)cap);
  const uint32_t tile = 1u << clamp_tile_log(cfg.slice_steps);
  const uint32_t nm_cap = (uint32_t)((k_cap + tile - 1) / tile);
  rz_status st;
  ctx->stamp_groups = 0;
  if ((st = ensure(ctx, ctx->tail_bits, (size_t)((k_cap + 63) / 64) * 8, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->mstatus, 2 * (size_t)kMaxSlices * 8, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->slots, ((size_t)k_cap / 2 + 1) * 8, s)) != RZ_OK) return st;
  uint32_t* ctrl = (uint32_t*)ctx->ctrl.ptr;
  const int stats = (flags & RZ_STATS) ? 1 : 0;
  if (stats) {
    if ((st = fill_u32(ctrl + kCtrlStatTail, 0u, 2, s)) != RZ_OK) return st;
    if ((st = fill_u32(ctrl + kCtrlStatOrphans, 0u, 2, s)) != RZ_OK) return st;
  }
  if ((st = prof_mark(ctx, flags, 0, s)) != RZ_OK) return st;
  if ((st = prof_mark(ctx, flags, 1, s)) != RZ_OK) return st;
  uint32_t* bits = (uint32_t*)ctx->tail_bits.ptr;
  uint64_t* mstatus = (uint64_t*)ctx->mstatus.ptr;
  uint64_t* slots = (uint64_t*)ctx->slots.ptr;
  const uint32_t bo = kMarkWaitBackoff, nc = (uint32_t)n_cap, kc = (uint32_t)k_cap;
  const int sorted = (flags & RZ_SORTED) ? 1 : 0, early = (flags & RZ_ANY_PAIRING) ? 1 : 0;
  const size_t lds = (size_t)L_cap * 4;
  bool launched = false;
#define RZ_SZ_CASE(W, I)                                                                                         \
  if (!launched && cfg.wg_size == W && cfg.items == I) {                                                         \
    hipLaunchKernelGGL((rz_mark_move_pair_sized<T, W, I>), dim3(nm_cap + ns_cap), dim3(W), lds, s, A, R, d_n, d_k, \
                       nc, kc, cap, nm_cap, bits, mstatus, slots, ctrl, stats, tile, bo, sorted, early);          \
    launched = true;                                                                                             \
  }
  RZ_SZ_CASE(256, 2) RZ_SZ_CASE(256, 4) RZ_SZ_CASE(256, 8) RZ_SZ_CASE(256, 16)
  RZ_SZ_CASE(512, 2) RZ_SZ_CASE(512, 4) RZ_SZ_CASE(512, 8) RZ_SZ_CASE(512, 16)
#undef RZ_SZ_CASE
  if (!launched) return RZ_EINVAL;
  RZ_HIP(hipGetLastError());
  if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
  return prof_mark(ctx, flags, 3, s);
}

rz_status launch_sized(rz_ctx* ctx, void* data, int64_t n_cap, uint32_t esz, const int32_t* R, int64_t k_cap,
                       const int32_t* d_n, const int32_t* d_k, uint32_t flags, const rz_tile_config& cfg,
                       hipStream_t s) {
  switch (esz) {
    case 4: return sized_impl<uint32_t>(ctx, (uint32_t*)data, n_cap, R, k_cap, d_n, d_k, flags, cfg, s);
    case 8: return sized_impl<uint64_t>(ctx, (uint64_t*)data, n_cap, R, k_cap, d_n, d_k, flags, cfg, s);
    case 16: return sized_impl<uint4>(ctx, (uint4*)data, n_cap, R, k_cap, d_n, d_k, flags, cfg, s);
    default: return RZ_EINVAL;
  }
}

}  // namespace rz
