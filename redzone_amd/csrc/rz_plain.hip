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
// rz_batch.hip with a single list.
//
// Contract violations (duplicates, out-of-range indices) never cause an
// out-of-bounds access: marks are taken only for indices inside the tail,
// moves only into holes below n-k, and pairing stops at the smaller orphan
// total (a mismatch sets an error bit in the control block).
#include "rz_redzone_dev.hpp"

namespace rz {

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
__global__ __launch_bounds__(256) void rz_pair(T* __restrict__ A, const int32_t* __restrict__ Rd, uint32_t sZ,
                                                uint32_t L, uint32_t ns, const uint64_t* __restrict__ counts,
                                                uint32_t* __restrict__ ctrl, int stats) {
  __shared__ PairShared<256, kMaxSlices> sh;
  pair_body<T, 256, kMaxSlices, false>(A, Rd, sZ, L, ns, counts, ctrl, stats, blockIdx.x, gridDim.x, sh);
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
    while (ld_relaxed(cnt) < target) {
      __builtin_amdgcn_s_sleep(2);
      if (++spins > (1u << 22)) {
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
  pair_body<T, WG, kMaxSlices, true>(A, Rd, sZ, L, ns, counts, ctrl, stats, blockIdx.x, G, psh);
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
    pair_body_lds<T, WG, true>(A, Rd, sZ, L, ns, counts, ctrl, stats, b - nm - ns, gridDim.x - nm - ns, dyn,
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
                            const rz_tile_config& cfg, bool coop, hipStream_t s) {
  const uint32_t sZ = (uint32_t)(n - k);
  uint32_t L, ns;
  slice_shape(k, cfg, kMaxSlices, &L, &ns);
  rz_status st;
  if ((st = ensure(ctx, ctx->tail_bits, (size_t)((k + 63) / 64) * 8, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->counts, (size_t)ns * 8, s)) != RZ_OK) return st;
  int32_t* Rd = R;
  if (flags & RZ_PRESERVE_REMOVALS) {
    if ((st = ensure(ctx, ctx->orphans, (size_t)k * 4, s)) != RZ_OK) return st;
    Rd = (int32_t*)ctx->orphans.ptr;
  }
  uint32_t* bits = (uint32_t*)ctx->tail_bits.ptr;
  uint64_t* counts = (uint64_t*)ctx->counts.ptr;
  uint32_t* ctrl = (uint32_t*)ctx->ctrl.ptr;
  const int stats = (flags & RZ_STATS) ? 1 : 0;
  if (stats && (st = fill_u32(ctrl + kCtrlStatTail, 0u, 2, s)) != RZ_OK) return st;
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
  const int64_t g1w = (k + 256 * 4 - 1) / (256 * 4);
  const dim3 g1((uint32_t)(g1w < (int64_t)ctx->cu_count * 8 ? g1w : ctx->cu_count * 8));
  if (stats)
    hipLaunchKernelGGL(rz_mark_tail<true>, g1, dim3(256), 0, s, R, (uint32_t)k, sZ, bits,
                       (uint64_t*)(ctrl + kCtrlStatTail));
  else if (flags & RZ_SORTED)
    hipLaunchKernelGGL(rz_mark_tail_sorted, dim3((uint32_t)((k + kSortedChunk - 1) / kSortedChunk)), dim3(256), 0,
                       s, R, (uint32_t)k, sZ, bits);
  else
    hipLaunchKernelGGL(rz_mark_tail<false>, g1, dim3(256), 0, s, R, (uint32_t)k, sZ, bits, (uint64_t*)nullptr);
  RZ_HIP(hipGetLastError());
  if ((st = prof_mark(ctx, flags, 1, s)) != RZ_OK) return st;
  if (!dispatch_move<T>(cfg.wg_size, cfg.items, dim3(ns), s, A, R, Rd, (uint32_t)k, sZ, L, bits, counts))
    return RZ_EINVAL;
  RZ_HIP(hipGetLastError());
  if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
  const dim3 g3((uint32_t)ctx->cu_count);
  hipLaunchKernelGGL(rz_pair<T>, g3, dim3(256), 0, s, A, Rd, sZ, L, ns, counts, ctrl, stats);
  RZ_HIP(hipGetLastError());
  return prof_mark(ctx, flags, 3, s);
}

rz_status launch_plain(rz_ctx* ctx, void* data, int64_t n, uint32_t esz, int32_t* R, int64_t k, uint32_t flags,
                       const rz_tile_config& cfg, hipStream_t s) {
  if (cfg.variant == 2) return launch_queue_single(ctx, data, n, esz, R, k, flags, cfg, s);
  const bool coop = cfg.variant == 1 || (flags & RZ_COOPERATIVE);
  switch (esz) {
    case 4: return plain_impl<uint32_t>(ctx, (uint32_t*)data, n, R, k, flags, cfg, coop, s);
    case 8: return plain_impl<uint64_t>(ctx, (uint64_t*)data, n, R, k, flags, cfg, coop, s);
    case 16: return plain_impl<uint4>(ctx, (uint4*)data, n, R, k, flags, cfg, coop, s);
    default: return RZ_EINVAL;
  }
}

}  // namespace rz
