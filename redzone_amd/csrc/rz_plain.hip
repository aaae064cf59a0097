// rz_plain.hip -- Redzone O(k) removal from one list on gfx950.
//
// The removal list R (k entries) pairs with the red zone Z = A[n-k, n): pair i
// moves Z[i] into hole R[i].  Three dependent steps, each a kernel on the
// caller's stream (the kernel boundary is the grid-wide barrier; a regular
// launch on MI355X costs ~2 us, a cooperative launch ~20 us, and an in-kernel
// barrier that fences at agent scope ~14 us at 256 workgroups):
//
//   1. rz_mark_tail   every R[i] >= n-k kills tail item R[i]-(n-k): one bit in
//                     a k-bit tail bitmap (relaxed atomicOr).  Replaces the
//                     sentinel write of redzone.hxx:698-707, so no payload
//                     value is reserved.
//   2. rz_move        R is cut into slices of L pairs, one workgroup each.  A
//                     workgroup loads its R, Z and bitmap words coalesced,
//                     classifies every pair with two 64-lane ballots, and ranks
//                     the orphans of its slice in i order with one LDS scan per
//                     step (redzone.hxx:714-822 does two block scans per step
//                     on 32-lane warps).  Moves go straight to their holes;
//                     dest orphans go to the slice's own part of R, source
//                     orphans to the slice's own part of Z.  The slice's
//                     bitmap words are cleared after use, so the bitmap is
//                     zero again for the next call without a memset.
//   3. rz_pair        every workgroup scans the per-slice orphan counts in LDS
//                     and pairs the t-th dest orphan with the t-th source
//                     orphan (redzone.hxx:834-934).  Orphans are ordered by
//                     slice then by i, so the result is the canonical one for
//                     every launch shape.
//
// Contract violations (duplicates, out-of-range indices) never cause an
// out-of-bounds access: marks are taken only for indices inside the tail,
// moves only into holes below n-k, and pairing stops at the smaller orphan
// total (a mismatch sets an error bit in the control block).
#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

// ---------------------------------------------------------------------------
// coherent element access (used where another workgroup of the same kernel
// reads the value: the cooperative single-launch variant)
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
// 1. mark
// ---------------------------------------------------------------------------
template <bool STATS>
__device__ __forceinline__ void mark_body(const int32_t* __restrict__ R, uint32_t k, uint32_t sZ,
                                          uint32_t* __restrict__ bits, uint64_t* __restrict__ stat_tail,
                                          uint32_t gtid, uint32_t stride) {
  constexpr int U = 4;
  uint32_t hits = 0;
  for (uint32_t i0 = gtid; i0 < k; i0 += stride * U) {
    uint32_t r[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * stride;
      r[u] = i < k ? (uint32_t)R[i] : sZ - 1u;  // sZ-1 is never inside the tail (k > 0 => sZ < n)
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

template <bool STATS>
__global__ __launch_bounds__(256) void rz_mark_tail(const int32_t* __restrict__ R, uint32_t k, uint32_t sZ,
                                                     uint32_t* __restrict__ bits, uint64_t* __restrict__ stat_tail) {
  mark_body<STATS>(R, k, sZ, bits, stat_tail, blockIdx.x * 256u + threadIdx.x, gridDim.x * 256u);
}

// ---------------------------------------------------------------------------
// 2. move and collect orphans
// ---------------------------------------------------------------------------
// Item classes.  A pair with a dead source and an invalid destination does nothing.
enum : uint32_t { kNone = 0, kMove = 1, kSrcOrphan = 2, kDstOrphan = 3 };

template <int WG, int IPT>
struct MoveShared {
  static constexpr int E = IPT * (WG / kWave);
  uint32_t off[E + 1];
};

// One slice [a, b) of pairs.  COH: bitmap read and orphans / counts written
// coherently for readers in other workgroups of the same kernel.
template <typename T, int WG, int IPT, bool COH>
__device__ __forceinline__ void move_slice(T* __restrict__ A, const int32_t* __restrict__ R, int32_t* __restrict__ Rd,
                                           uint32_t k, uint32_t sZ, uint32_t L, uint32_t slice,
                                           uint32_t* __restrict__ bits, uint64_t* __restrict__ counts,
                                           MoveShared<WG, IPT>& sh) {
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
        wv[j] = ld_x<COH>(&bits[p >> 5]);
      } else {
        d[j] = sZ;            // invalid destination
        wv[j] = 0xffffffffu;  // dead source
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
#pragma unroll
    for (int j = 0; j < IPT; ++j) {
      const uint32_t cls = info[j] & 0xffu;
      if (cls == kNone) continue;
      if (cls == kMove) {
        A[d[j]] = z[j];
      } else {
        const uint32_t off = s_off[j * NW + wave];
        if (cls == kSrcOrphan) {
          st_x<COH>(&Z[a + carry_s + (off & 0xffffu) + ((info[j] >> 8) & 0xffu)], z[j]);
        } else {
          st_x<COH>((uint32_t*)&Rd[a + carry_d + (off >> 16) + (info[j] >> 16)], d[j]);
        }
      }
    }
    const uint32_t tot = s_off[E];
    carry_s += tot & 0xffffu;
    carry_d += tot >> 16;
    __syncthreads();
  }
  // leave the bitmap zero for the next call (slices are word aligned)
  for (uint32_t w = (a >> 5) + tid; w < ((b + 31u) >> 5); w += WG) bits[w] = 0u;
  if (tid == 0) st_x<COH>(&counts[slice], (uint64_t)carry_d | ((uint64_t)carry_s << 32));
}

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_move(T* __restrict__ A, const int32_t* __restrict__ R,
                                               int32_t* __restrict__ Rd, uint32_t k, uint32_t sZ, uint32_t L,
                                               uint32_t* __restrict__ bits, uint64_t* __restrict__ counts) {
  __shared__ MoveShared<WG, IPT> sh;
  move_slice<T, WG, IPT, false>(A, R, Rd, k, sZ, L, blockIdx.x, bits, counts, sh);
}

// ---------------------------------------------------------------------------
// 3. pair orphans
// ---------------------------------------------------------------------------
template <int WG>
struct PairShared {
  uint64_t pre[kMaxSlices + 1];
  uint64_t tmp[WG / kWave + 1];
};

// Exclusive prefix of the per-slice counts into LDS, then pairs t = start,
// start+stride, ...  Dest orphans are the low 32 bits of each count word.
template <typename T, int WG, bool COH>
__device__ __forceinline__ void pair_body(T* __restrict__ A, const int32_t* __restrict__ Rd, uint32_t sZ,
                                          uint32_t L, uint32_t ns, const uint64_t* __restrict__ counts,
                                          uint32_t* __restrict__ ctrl, int stats, uint32_t block,
                                          uint32_t nblocks, PairShared<WG>& sh) {
  constexpr int PER = kMaxSlices / WG;
  uint64_t* s_pre = sh.pre;
  const int tid = threadIdx.x;
  uint64_t loc[PER];
  uint64_t sum = 0;
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const uint32_t s = tid * PER + q;
    loc[q] = s < ns ? ld_x<COH>(&counts[s]) : 0ull;
    sum += loc[q];
  }
  uint64_t total;
  uint64_t ex = block_excl_scan<WG>(sum, sh.tmp, &total);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const uint32_t s = tid * PER + q;
    if (s < ns) s_pre[s] = ex;
    ex += loc[q];
  }
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
    const uint32_t dst = ld_x<COH>((const uint32_t*)&Rd[sd * L + (t - (uint32_t)s_pre[sd])]);
    A[dst] = ld_x<COH>(&Z[ss * L + (t - (uint32_t)(s_pre[ss] >> 32))]);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void rz_pair(T* __restrict__ A, const int32_t* __restrict__ Rd, uint32_t sZ,
                                                uint32_t L, uint32_t ns, const uint64_t* __restrict__ counts,
                                                uint32_t* __restrict__ ctrl, int stats) {
  __shared__ PairShared<256> sh;
  pair_body<T, 256, false>(A, Rd, sZ, L, ns, counts, ctrl, stats, blockIdx.x, gridDim.x, sh);
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
  __shared__ PairShared<WG> psh;
  const uint64_t base = ld_relaxed((const uint64_t*)&ctrl[kCtrlBarEpoch]);
  const uint32_t G = gridDim.x;
  if (stats)
    mark_body<true>(R, k, sZ, bits, (uint64_t*)&ctrl[kCtrlStatTail], blockIdx.x * WG + threadIdx.x, G * WG);
  else
    mark_body<false>(R, k, sZ, bits, nullptr, blockIdx.x * WG + threadIdx.x, G * WG);
  grid_barrier(ctrl, base + G);
  // every workgroup has read `base` before arriving: publish the next launch's base
  if (blockIdx.x == 0 && threadIdx.x == 0) st_relaxed((uint64_t*)&ctrl[kCtrlBarEpoch], base + 2ull * G);
  for (uint32_t s = blockIdx.x; s < ns; s += G) move_slice<T, WG, IPT, true>(A, R, Rd, k, sZ, L, s, bits, counts, msh);
  grid_barrier(ctrl, base + 2ull * G);
  pair_body<T, WG, true>(A, Rd, sZ, L, ns, counts, ctrl, stats, blockIdx.x, G, psh);
}

// ---------------------------------------------------------------------------
// launch
// ---------------------------------------------------------------------------
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

// Slice length: a multiple of one step (wg * items), with at most max_slices slices.
static void slice_shape(int64_t k, const rz_tile_config& cfg, uint32_t* L, uint32_t* ns) {
  const int64_t step = (int64_t)cfg.wg_size * cfg.items;
  int64_t cap = cfg.max_slices < 1 ? 1 : cfg.max_slices;
  if (cap > kMaxSlices) cap = kMaxSlices;
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
  slice_shape(k, cfg, &L, &ns);
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
  if (coop) {
    if ((st = prof_mark(ctx, flags, 1, s)) != RZ_OK) return st;
    st = dispatch_coop<T>(cfg.wg_size, cfg.items, ctx, A, R, Rd, (uint32_t)k, sZ, L, ns, bits, counts, ctrl, stats, s);
    if (st != RZ_OK) return st;
    if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
    return prof_mark(ctx, flags, 3, s);
  }
  const int64_t g1w = (k + 256 * 4 - 1) / (256 * 4);
  const dim3 g1((uint32_t)(g1w < (int64_t)ctx->cu_count * 8 ? g1w : ctx->cu_count * 8));
  if (stats)
    hipLaunchKernelGGL(rz_mark_tail<true>, g1, dim3(256), 0, s, R, (uint32_t)k, sZ, bits,
                       (uint64_t*)(ctrl + kCtrlStatTail));
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
  const bool coop = cfg.variant == 1 || (flags & RZ_COOPERATIVE);
  switch (esz) {
    case 4: return plain_impl<uint32_t>(ctx, (uint32_t*)data, n, R, k, flags, cfg, coop, s);
    case 8: return plain_impl<uint64_t>(ctx, (uint64_t*)data, n, R, k, flags, cfg, coop, s);
    case 16: return plain_impl<uint4>(ctx, (uint4*)data, n, R, k, flags, cfg, coop, s);
    default: return RZ_EINVAL;
  }
}

}  // namespace rz
