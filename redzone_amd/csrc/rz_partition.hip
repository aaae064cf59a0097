// rz_partition.hip -- O(n) stable in-place partition on gfx950, the comparator
// and dispatch alternative of the Redzone path, plus input validation.
//
// The reference's O(n) paths mark each removal with a -1 sentinel
// (remove_phase1, DataStruct.hxx:11-18) and compact items != -1 in one block
// (StreamCompact_OneBlock, redzone.hxx:526-603) or with thrust::remove
// (Thrust_remove.hxx:15-17).  Here:
//   rz_mark_all   sets one bit per removal in an n-bit bitmap (payload
//                 agnostic) and bumps the look-back epoch; lists of 2^23
//                 items and more with 2^20 removals and more group the
//                 marks by bitmap region first (bucket_marks, rz_dense.hip)
//                 and write every bitmap word whole;
//   rz_compact    one pass over all n items in tiles of WG*IPT, tile id =
//                 workgroup id (workgroups are dispatched in id order).  Each
//                 tile ranks its kept items with wave ballots and one LDS scan,
//                 publishes its count, finds its output offset by decoupled
//                 look-back over the predecessors' status words, and writes
//                 its kept items in order.  A tile publishes only after all of
//                 its loads have returned, and a tile writes only below the
//                 end of its own input, so the in-place pass never overwrites
//                 an item before it was read.  It clears the bitmap words it
//                 read, so the bitmap is zero again for the next call.
//                 The 16-byte aligned form (rz_compact_vec) publishes a
//                 tile's kept count as soon as it has the tile's bitmap words
//                 (the count needs no item), so the look-back never waits for
//                 loads; the in-place guard is separate: a tile sets its
//                 loaded flag (and its group of 64 tiles a group word) once
//                 its items are in registers, and a tile writes only after
//                 every tile overlapping its output range has loaded.  It
//                 stages the kept items in LDS and writes them as whole
//                 vectors.
#include <atomic>

#include "rz_compact_dev.hpp"
#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

__global__ __launch_bounds__(256) void rz_mark_all(const int32_t* __restrict__ R, uint32_t k, uint32_t n,
                                                    uint32_t* __restrict__ bits, uint32_t* __restrict__ ctrl) {
  const uint32_t stride = gridDim.x * 256u;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < k; i += stride) {
    const uint32_t r = (uint32_t)R[i];
    if (r < n) or_relaxed(&bits[r >> 5], 1u << (r & 31));
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) add_relaxed(&ctrl[kCtrlEpoch], 1u);
}

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_compact(T* __restrict__ A, uint32_t n, uint32_t* __restrict__ bits,
                                                  uint64_t* __restrict__ status, uint32_t* __restrict__ ctrl) {
  const uint32_t epoch = ld_relaxed(&ctrl[kCtrlEpoch]) & 0x3fffffffu;
  compact_tile<T, WG, IPT, false>(A, n, bits, status, blockIdx.x, epoch, ctrl);
}

// ---------------------------------------------------------------------------
// Vectorised form: every thread loads whole 16-byte vectors (V = 16 / sizeof(T)
// items), the tile's bitmap words are read once into LDS (and cleared at
// once), and a lane's rank among the kept items of a wave comes from three
// bit-sliced ballots of its per-vector keep count.  The kept items are staged
// in LDS in tile order while wave 0 looks back, then written out as aligned
// 16-byte vectors (single items only at the two ends).  Needs a 16-byte
// aligned list.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t keep_count_prefix(uint32_t c, uint32_t* total) {
  // c in [0, 4]: three bit planes
  const uint64_t b0 = ballot(c & 1u), b1 = ballot(c & 2u), b2 = ballot(c & 4u);
  *total = (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
  return lane_rank(b0) + 2u * lane_rank(b1) + 4u * lane_rank(b2);
}

// item q of a 16-byte vector, and a vector of V items
template <typename T, int Q> __device__ __forceinline__ T vec_item(const uint4& v) {
  if constexpr (sizeof(T) == 4) {
    return Q == 0 ? v.x : Q == 1 ? v.y : Q == 2 ? v.z : v.w;
  } else if constexpr (sizeof(T) == 8) {
    return Q == 0 ? ((uint64_t)v.y << 32 | v.x) : ((uint64_t)v.w << 32 | v.z);
  } else {
    return v;
  }
}
template <typename T, int Q> __device__ __forceinline__ void vec_set(uint4& v, const T& e) {
  if constexpr (sizeof(T) == 4) {
    if (Q == 0) v.x = e; else if (Q == 1) v.y = e; else if (Q == 2) v.z = e; else v.w = e;
  } else if constexpr (sizeof(T) == 8) {
    if (Q == 0) { v.x = (uint32_t)e; v.y = (uint32_t)(e >> 32); } else { v.z = (uint32_t)e; v.w = (uint32_t)(e >> 32); }
  } else {
    v = e;
  }
}

template <typename T, int V, int Q = 0>
__device__ __forceinline__ void load_partial(const T* __restrict__ A, uint32_t p0, uint32_t n, uint4& v) {
  if constexpr (Q < V) {
    if (p0 + Q < n) vec_set<T, Q>(v, A[p0 + Q]);
    load_partial<T, V, Q + 1>(A, p0, n, v);
  }
}

template <typename T, int V, int Q = 0>
__device__ __forceinline__ void stage_kept(T* __restrict__ stage, const uint4& v, uint32_t km, uint32_t pos) {
  if constexpr (Q < V) {
    if ((km >> Q) & 1u) stage[pos++] = vec_item<T, Q>(v);
    stage_kept<T, V, Q + 1>(stage, v, km, pos);
  }
}

template <typename T, int V, int Q = 0>
__device__ __forceinline__ void gather_vec(const T* __restrict__ stage, uint32_t o, uint4& v) {
  if constexpr (Q < V) {
    vec_set<T, Q>(v, stage[o + Q]);
    gather_vec<T, V, Q + 1>(stage, o, v);
  }
}

template <typename T, int WG, int VPT>
struct CompactShared {
  static constexpr int NW = WG / kWave;
  static constexpr uint32_t TILE = WG * VPT * (16 / sizeof(T));
  T stage[TILE];
  uint32_t bits[TILE / 32];
  uint32_t off[VPT * NW + 1];
  uint32_t excl;
};

template <typename T, int WG, int VPT>
__global__ __launch_bounds__(WG) void rz_compact_vec(T* __restrict__ A, uint32_t n, uint32_t* __restrict__ bits,
                                                      uint64_t* __restrict__ status, uint32_t* __restrict__ loaded,
                                                      uint32_t* __restrict__ gcnt, uint32_t* __restrict__ gdone,
                                                      uint32_t ntiles, uint32_t* __restrict__ ctrl) {
  constexpr int V = 16 / sizeof(T);
  constexpr int NW = WG / kWave;
  constexpr int E = VPT * NW;
  constexpr int PER = (E + kWave - 1) / kWave;
  constexpr uint32_t TILE = WG * VPT * V;
  constexpr int WORDS = TILE / 32;
  static_assert(TILE % 32 == 0, "tiles are bitmap-word aligned");
  __shared__ CompactShared<T, WG, VPT> sh;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint32_t t = blockIdx.x;
  const uint32_t base = t * TILE;
  const uint32_t epoch = ld_relaxed(&ctrl[kCtrlEpoch]) & 0x3fffffffu;
  const uint32_t wend = (n + 31u) >> 5;

  // bitmap words of the tile -> LDS, and leave them zero for the next call
  for (int w = tid; w < WORDS; w += WG) {
    const uint32_t gw = (base >> 5) + w;
    uint32_t v = 0xffffffffu;
    if (gw < wend) {
      v = bits[gw];
      bits[gw] = 0u;
    }
    sh.bits[w] = v;
  }
  uint4 x[VPT];
  auto load_items = [&]() {
#pragma unroll
    for (int j = 0; j < VPT; ++j) {
      const uint32_t p0 = base + (uint32_t)(j * WG + tid) * V;
      if (p0 + V <= n) {
        x[j] = *(const uint4*)(A + p0);
      } else {
        x[j] = make_uint4(0u, 0u, 0u, 0u);
        load_partial<T, V>(A, p0, n, x[j]);
      }
    }
  };
  // wave 0 loads its items after it has published the tile's count, so that
  // its status store does not wait behind them (loads and stores retire in order)
  if (wave != 0) load_items();
  __syncthreads();
  // the tile's kept count depends on the bitmap alone
  uint32_t keep[VPT], pre[VPT];
#pragma unroll
  for (int j = 0; j < VPT; ++j) {
    const uint32_t r0 = (uint32_t)(j * WG + tid) * V;  // offset in the tile (V items share one bitmap word)
    const uint32_t p0 = base + r0;
    uint32_t km = (~(sh.bits[r0 >> 5] >> (r0 & 31))) & ((1u << V) - 1u);
    if (p0 + V > n) km &= p0 < n ? (1u << (n - p0)) - 1u : 0u;
    keep[j] = km;
    uint32_t tot;
    pre[j] = keep_count_prefix((uint32_t)__popc(km), &tot);
    if (lane == 0) sh.off[j * NW + wave] = tot;
  }
  __syncthreads();
  if (wave == 0) {
    uint32_t loc[PER];
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int idx = lane * PER + q;
      loc[q] = idx < E ? sh.off[idx] : 0u;
      sum += loc[q];
    }
    const uint32_t xs = wave_incl_scan(sum);
    uint32_t ex = xs - sum;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int idx = lane * PER + q;
      if (idx < E) sh.off[idx] = ex;
      ex += loc[q];
    }
    if (lane == kWave - 1) sh.off[E] = xs;
    const uint32_t agg = __shfl(xs, kWave - 1, kWave);
    // the count goes out before any item is read: the offsets of later
    // tiles never wait for loads
    if (lane == 0) st_relaxed(&status[t], pack_status(t == 0 ? kFlagIncl : kFlagAgg, epoch, agg));
    load_items();
  }
  wait_all_memory();  // every item load of this tile has returned
  __syncthreads();
  if (tid == 0) {
    // this tile's input may now be overwritten: its flag, and its group's
    // word once all 64 tiles of the group (fewer in the last) have loaded.
    // The counters and the words live in separate arrays: words polled next
    // to counters being incremented slowed the pass by half.
    st_relaxed(&loaded[t], epoch);
    const uint32_t g = t >> 6, members = min(64u, ntiles - (g << 6));
    if (add_relaxed(&gcnt[g], 1u) == members - 1u) {
      st_relaxed(&gcnt[g], 0u);
      st_relaxed(&gdone[g], epoch);
    }
  }
  // kept items in tile order into LDS (wave 0 stages its items, then looks back)
#pragma unroll
  for (int j = 0; j < VPT; ++j) stage_kept<T, V>(sh.stage, x[j], keep[j], sh.off[j * NW + wave] + pre[j]);
  if (wave == 0) {
    const uint32_t agg = sh.off[E];
    uint32_t excl = 0;
    if (t != 0) {
      // The nearest predecessor's word alone first (one load per poll, pause
      // doubling to 16 x 64 cycles), then windows of 64 back to an inclusive
      // prefix: polling whole windows from every waiting tile costs more
      // fabric bandwidth than the moves.
      uint32_t pause = 1, spins = 0;
      while (true) {
        const uint64_t st = ld_relaxed(&status[t - 1]);
        if ((((uint32_t)(st >> 32) & 0x3fffffffu) == epoch && (st >> 62) != 0u) || ++spins > spin_cap(ctrl, 1u << 22)) break;
        for (uint32_t i = 0; i < pause; ++i) __builtin_amdgcn_s_sleep(1);
        if (pause < 16u) pause <<= 1;
      }
      int64_t jb = (int64_t)t - 1;
      spins = 0;
      pause = 1;
      while (true) {
        const int64_t q = jb - lane;
        uint64_t st = q >= 0 ? ld_relaxed(&status[q]) : pack_status(kFlagIncl, epoch, 0u);
        uint32_t flag = (((uint32_t)(st >> 32) & 0x3fffffffu) == epoch) ? (uint32_t)(st >> 62) : 0u;
        if (q < 0) flag = (uint32_t)kFlagIncl;
        const uint64_t incl_m = ballot(flag == kFlagIncl);
        const uint64_t none_m = ballot(flag == 0u);
        const uint64_t upto = incl_m ? ((incl_m & (~incl_m + 1)) << 1) - 1 : ~0ull;
        if (none_m & upto) {
          if (++spins > spin_cap(ctrl, 1u << 22)) {
            if (lane == 0) or_relaxed(&ctrl[kCtrlError], kErrLookbackTimeout);
            break;
          }
          for (uint32_t i = 0; i < pause; ++i) __builtin_amdgcn_s_sleep(1);
          if (pause < 16u) pause <<= 1;
          continue;
        }
        const uint32_t v = ((upto >> lane) & 1ull) ? (uint32_t)st : 0u;
        excl += wave_sum(v);
        if (incl_m) break;
        jb -= kWave;
      }
      if (lane == 0) st_relaxed(&status[t], pack_status(kFlagIncl, epoch, excl + agg));
      // The pass is in place: every tile whose input overlaps [excl, base)
      // must have loaded before this tile writes.  Those are tiles [u, t):
      // whole groups of 64 by their word, the two ends tile by tile.
      const uint32_t u = excl / TILE;
      const uint32_t g0 = (u + 63u) >> 6, g1 = t >> 6;
      const bool groups = g0 < g1;
      const uint32_t a_end = groups ? (g0 << 6) : t, b_beg = groups ? (g1 << 6) : t;
      const uint32_t na = a_end - u, ng = groups ? g1 - g0 : 0u, nb = t - b_beg;
      const uint32_t total = na + ng + nb;
      spins = 0;
      pause = 1;
      for (uint32_t c0 = 0; c0 < total;) {
        const uint32_t c = c0 + lane;
        bool ok = true;
        if (c < na)
          ok = ld_relaxed(&loaded[u + c]) == epoch;
        else if (c < na + ng)
          ok = ld_relaxed(&gdone[g0 + (c - na)]) == epoch;
        else if (c < total)
          ok = ld_relaxed(&loaded[b_beg + (c - na - ng)]) == epoch;
        if (ballot(!ok)) {
          if (++spins > spin_cap(ctrl, 1u << 22)) {
            if (lane == 0) or_relaxed(&ctrl[kCtrlError], kErrLookbackTimeout);
            break;
          }
          for (uint32_t i = 0; i < pause; ++i) __builtin_amdgcn_s_sleep(1);
          if (pause < 16u) pause <<= 1;
          continue;
        }
        c0 += kWave;
      }
    }
    if (lane == 0) sh.excl = excl;
  }
  __syncthreads();
  // A[excl, excl + agg) = stage[0, agg): whole aligned vectors, single items at the ends
  const uint32_t lo = sh.excl, hi = lo + sh.off[E];
  const uint32_t va = (lo + V - 1) / V * V, vb = hi / V * V;
  if (va >= vb) {
    for (uint32_t p = lo + tid; p < hi; p += WG) A[p] = sh.stage[p - lo];
  } else {
    for (uint32_t p = lo + tid; p < va; p += WG) A[p] = sh.stage[p - lo];
    for (uint32_t p = vb + tid; p < hi; p += WG) A[p] = sh.stage[p - lo];
    for (uint32_t p = va + (uint32_t)tid * V; p < vb; p += WG * V) {
      uint4 v;
      gather_vec<T, V>(sh.stage, p - lo, v);
      *(uint4*)(A + p) = v;
    }
  }
}

// ---------------------------------------------------------------------------
// validation: range and duplicates (uses the n-bit bitmap, cleared afterwards)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void rz_validate_mark(const int32_t* __restrict__ R, uint32_t k, uint32_t n,
                                                         uint32_t* __restrict__ bits, uint32_t* __restrict__ flag,
                                                         int check_sorted) {
  uint32_t bad = 0;
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < k; i += gridDim.x * 256u) {
    const uint32_t r = (uint32_t)R[i];
    if (check_sorted && i + 1 < k && (uint32_t)R[i + 1] <= r) bad |= 4u;
    if (r >= n) { bad |= 1u; continue; }
    const uint32_t bit = 1u << (r & 31);
    if (or_relaxed(&bits[r >> 5], bit) & bit) bad |= 2u;
  }
  if (bad) or_relaxed(flag, bad);
}

__global__ __launch_bounds__(256) void rz_validate_clear(const int32_t* __restrict__ R, uint32_t k, uint32_t n,
                                                          uint32_t* __restrict__ bits) {
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < k; i += gridDim.x * 256u) {
    const uint32_t r = (uint32_t)R[i];
    if (r < n) bits[r >> 5] = 0u;
  }
}

rz_status launch_validate(rz_ctx* ctx, const int32_t* R, int64_t n, int64_t k, uint32_t flags, hipStream_t s) {
  rz_status st;
  if ((st = ensure(ctx, ctx->list_bits, (size_t)((n + 63) / 64) * 8, s)) != RZ_OK) return st;
  uint32_t* ctrl = (uint32_t*)ctx->ctrl.ptr;
  if ((st = fill_u32(ctrl + kCtrlValidate, 0u, 1, s)) != RZ_OK) return st;
  const int64_t gw = (k + 255) / 256;
  const dim3 g((uint32_t)(gw < (int64_t)ctx->cu_count * 8 ? (gw > 0 ? gw : 1) : ctx->cu_count * 8));
  hipLaunchKernelGGL(rz_validate_mark, g, dim3(256), 0, s, R, (uint32_t)k, (uint32_t)n,
                     (uint32_t*)ctx->list_bits.ptr, ctrl + kCtrlValidate, (flags & RZ_SORTED) ? 1 : 0);
  hipLaunchKernelGGL(rz_validate_clear, g, dim3(256), 0, s, R, (uint32_t)k, (uint32_t)n,
                     (uint32_t*)ctx->list_bits.ptr);
  RZ_HIP(hipGetLastError());
  RZ_HIP(hipMemcpyAsync(ctx->ctrl_host + kCtrlValidate, ctrl + kCtrlValidate, 4, hipMemcpyDeviceToHost, s));
  RZ_HIP(hipStreamSynchronize(s));
  const uint32_t f = ctx->ctrl_host[kCtrlValidate];
  if (f & 1u) return RZ_ERANGE;
  if (f & 2u) return RZ_EDUP;
  if (f & 4u) return RZ_EINVAL;  // RZ_SORTED given for a list that is not ascending
  return RZ_OK;
}

// ---------------------------------------------------------------------------
// launch
// ---------------------------------------------------------------------------
constexpr int kPartWG = 256, kPartIPT = 16;
// ---------------------------------------------------------------------------
// Bucketed marks, two kernels: the removals r in a window [base, base + m)
// set bit r - base of an m-bit bitmap that is zero on entry (the partition's
// n-bit bitmap: base 0, m = n; the dense form's tail bitmap: base n - k,
// m = k).  Bins of 2^blog positions (2^15 to 2^19, so that a window has 256 to
// 1,024 bins).
// rz_part_append: a workgroup takes 1,024 x IPT removals, counts them per bin in
// LDS, reserves room in each bin's entry list (2^blog entries, a bin's
// positions) with one atomic per bin, sorts them by bin in LDS and writes them
// out bin by bin, so consecutive threads write consecutive words.
// rz_part_bits: G workgroups per bin (G = 1 from 256 bins up) set the bin's
// bits in an LDS bitmap and store its words into the bitmap (OR-ed when
// G > 1); the last of a bin's workgroups puts the bin's count back to zero.
// Optionally workgroup 0 starts the compaction's epoch, and the window's
// removal count is summed into a stat word (zeroed by the append kernel).
// No scans and no epoch launch: two launches where a count / scan / entries /
// bits pipeline took five.
// ---------------------------------------------------------------------------
constexpr uint32_t kPartMinBinLog = 15, kPartMaxBinLog = 19;  // 4 to 64 KiB LDS bitmaps
constexpr uint32_t kPartMaxBins = 1024;                       // lists of up to 2^29 items
constexpr uint32_t kAppWG = 1024;
// removals per thread: the bigger the stage (dynamic LDS, 4 bytes a removal),
// the longer each bin's runs of consecutive words and the fewer count atomics
constexpr int kAppIPT = 16, kAppIPTSmall = 8;

template <int IPT>
__global__ __launch_bounds__(1024) void rz_part_append(const int32_t* __restrict__ R, uint32_t k, uint32_t base,
                                                        uint32_t m, uint32_t blog, uint32_t nbins,
                                                        uint32_t* __restrict__ ent, uint32_t* __restrict__ cur,
                                                        uint64_t* __restrict__ stat) {
  constexpr uint32_t kAppS = kAppWG * IPT;
  __shared__ uint32_t s_cnt[kPartMaxBins], s_lbase[kPartMaxBins], s_gbase[kPartMaxBins];
  __shared__ uint32_t s_tmp[kAppWG / kWave + 1];
  extern __shared__ __align__(16) unsigned char dyn[];
  uint32_t* s_stage = (uint32_t*)dyn;
  const uint32_t tid = threadIdx.x;
  const uint32_t a = blockIdx.x * kAppS, b = min(a + kAppS, k);
  for (uint32_t w = tid; w < nbins; w += kAppWG) s_cnt[w] = 0u;
  if (stat && blockIdx.x == 0 && tid == 0) st_relaxed(stat, (uint64_t)0);
  __syncthreads();
  uint32_t r[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {  // all loads first (a use inside this loop would wait for each one)
    const uint32_t i = a + j * kAppWG + tid;
    r[j] = i < b ? (uint32_t)R[i] : 0xffffffffu;
  }
  // positions in the window; 0xffffffff (>= m) for none (0xffffffff - base >= 2^31 > m)
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t p = r[j] - base;
    r[j] = p < m ? p : 0xffffffffu;
  }
  // a wave whose removals share one bin (a run) counts them with one atomic
#pragma unroll
  for (int j = 0; j < IPT; ++j) lds_bin_count(s_cnt, r[j] >> blog, r[j] < m);
  __syncthreads();
  // exclusive prefix over the bins (one bin per thread), then room in each bin's list
  uint32_t total;
  const uint32_t c = tid < nbins ? s_cnt[tid] : 0u;
  const uint32_t ex = block_excl_scan<kAppWG>(c, s_tmp, &total);
  // the reservation's atomics stay in flight while the removals are staged:
  // only the write-out below needs their results
  uint32_t gb = 0u;
  if (tid < nbins) {
    s_lbase[tid] = ex;
    if (c) gb = add_relaxed(&cur[tid], c);
    s_cnt[tid] = 0u;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t bz = r[j] >> blog;
    const uint32_t rank = lds_bin_add(s_cnt, bz, r[j] < m);
    if (r[j] < m) s_stage[s_lbase[bz] + rank] = r[j];
  }
  if (tid < nbins) s_gbase[tid] = gb;
  __syncthreads();
  for (uint32_t t = tid; t < total; t += kAppWG) {
    const uint32_t v = s_stage[t];
    const uint32_t bz = v >> blog;
    const uint32_t g = s_gbase[bz] + (t - s_lbase[bz]);
    if (g < (1u << blog)) ent[((size_t)bz << blog) + g] = v;  // more only with duplicate removals
  }
}

__global__ __launch_bounds__(1024) void rz_part_bits(const uint32_t* __restrict__ ent, uint32_t* __restrict__ cur,
                                                      uint32_t* __restrict__ bdone, uint32_t m, uint32_t blog,
                                                      uint32_t G, uint32_t* __restrict__ bits,
                                                      uint32_t* __restrict__ ctrl, uint64_t* __restrict__ stat) {
  const uint32_t W = 1u << (blog - 5), cap = 1u << blog;
  extern __shared__ __align__(16) unsigned char dyn[];
  uint32_t* lb = (uint32_t*)dyn;
  const uint32_t bin = blockIdx.x / G, g = blockIdx.x % G, tid = threadIdx.x;
  for (uint32_t w = tid; w < W; w += 1024u) lb[w] = 0u;
  const uint32_t c = min(cur[bin], cap);
  __syncthreads();
  const uint32_t share = ((c + G - 1) / G + 3u) & ~3u;  // a multiple of 4: 16-byte vector loads
  const uint32_t s = min(g * share, c), e = min(s + share, c);
  const uint32_t* be = ent + ((size_t)bin << blog);  // 16-byte aligned: bins start at multiples of 2^15
  const uint32_t z0 = bin << blog;
  // Each thread takes U consecutive entries.  A run of consecutive positions
  // (clustered removals) then puts the lanes of one LDS atomic U positions
  // apart, so at most two lanes share a bitmap word, where consecutive entries
  // per lane would put 32 lanes on one word.
  constexpr int U = 16;
  for (uint32_t q0 = s + tid * U; q0 < e; q0 += 1024u * U) {
    uint32_t r[U];
    if (q0 + U <= e) {  // all U entries: U/4 vector loads, issued together (no branch between them)
      uint4 x[U / 4];
#pragma unroll
      for (int v = 0; v < U / 4; ++v) x[v] = *(const uint4*)(be + q0 + 4u * v);
#pragma unroll
      for (int v = 0; v < U / 4; ++v) {
        r[4 * v] = x[v].x - z0;
        r[4 * v + 1] = x[v].y - z0;
        r[4 * v + 2] = x[v].z - z0;
        r[4 * v + 3] = x[v].w - z0;
      }
    } else {  // the end of the bin's share
#pragma unroll
      for (int i = 0; i < U; ++i) r[i] = q0 + i < e ? be[q0 + i] - z0 : 0xffffffffu;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (r[u] < cap) atomicOr(&lb[r[u] >> 5], 1u << (r[u] & 31));
  }
  __syncthreads();
  const uint32_t nw = (uint32_t)(((uint64_t)m + 31u) >> 5);
  for (uint32_t w = tid; w < W; w += 1024u) {
    const uint32_t gw = (z0 >> 5) + w;
    if (gw < nw && lb[w]) {
      if (G == 1) bits[gw] = lb[w];
      else or_relaxed(&bits[gw], lb[w]);
    }
  }
  if (tid == 0) {
    // every workgroup of the bin has read its count: the last puts it back to zero
    if (add_relaxed(&bdone[bin], 1u) == G - 1u) {
      st_relaxed(&cur[bin], 0u);
      st_relaxed(&bdone[bin], 0u);
    }
    if (stat && g == 0 && c) atomicAdd((unsigned long long*)stat, (unsigned long long)c);
    if (ctrl && blockIdx.x == 0) add_relaxed(&ctrl[kCtrlEpoch], 1u);
  }
}

bool bucket_marks_fit(int64_t m) { return m > 0 && m <= ((int64_t)kPartMaxBins << kPartMaxBinLog); }

rz_status bucket_marks(rz_ctx* ctx, const int32_t* R, int64_t k, uint32_t base, int64_t m, uint32_t* bits,
                       uint32_t* epoch_ctrl, uint64_t* stat, hipStream_t s) {
  if (!bucket_marks_fit(m) || k < 1 || k > INT32_MAX) return RZ_EINVAL;
  uint32_t blog = kPartMinBinLog;
  while (blog < kPartMaxBinLog && (m >> blog) > 512) ++blog;
  const uint32_t nbins = (uint32_t)((m + (1ll << blog) - 1) >> blog);
  rz_status st;
  if ((st = ensure(ctx, ctx->pent, (size_t)nbins << (blog + 2), s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->pcur, (size_t)nbins * 8, s)) != RZ_OK) return st;  // counts, then done counters
  constexpr size_t lds_max = (size_t)4 << (kPartMaxBinLog - 5);
  const size_t lds_bits = (size_t)4 << (blog - 5);
  // the big stage where the device gives a workgroup its LDS (plus about 12 KiB static)
  constexpr size_t lds_app = (size_t)4 * kAppWG * kAppIPT, lds_app_small = (size_t)4 * kAppWG * kAppIPTSmall;
  const bool big = lds_app + 16384 <= (size_t)ctx->lds_max;
  static std::atomic<uint64_t> attr_mask{0};
  const int dev = ctx->device;
  if (dev < 0 || dev >= 64 || !(attr_mask.load(std::memory_order_relaxed) & (1ull << dev))) {
    RZ_HIP(hipFuncSetAttribute((const void*)rz_part_bits, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_max));
    if (big)
      RZ_HIP(hipFuncSetAttribute((const void*)rz_part_append<kAppIPT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds_app));
    if (dev >= 0 && dev < 64) attr_mask.fetch_or(1ull << dev, std::memory_order_relaxed);
  }
  uint32_t* cur = (uint32_t*)ctx->pcur.ptr;
  uint32_t* ent = (uint32_t*)ctx->pent.ptr;
  if (big)
    hipLaunchKernelGGL(rz_part_append<kAppIPT>, dim3((uint32_t)((k + kAppWG * kAppIPT - 1) / (kAppWG * kAppIPT))),
                       dim3(kAppWG), lds_app, s, R, (uint32_t)k, base, (uint32_t)m, blog, nbins, ent, cur, stat);
  else
    hipLaunchKernelGGL(rz_part_append<kAppIPTSmall>,
                       dim3((uint32_t)((k + kAppWG * kAppIPTSmall - 1) / (kAppWG * kAppIPTSmall))), dim3(kAppWG),
                       lds_app_small, s, R, (uint32_t)k, base, (uint32_t)m, blog, nbins, ent, cur, stat);
  uint32_t G = ((uint32_t)ctx->cu_count + nbins - 1) / nbins;  // about one workgroup per CU
  G = G < 1 ? 1 : (G > 16 ? 16 : G);
  hipLaunchKernelGGL(rz_part_bits, dim3(nbins * G), dim3(1024), lds_bits, s, (const uint32_t*)ent, cur, cur + nbins,
                     (uint32_t)m, blog, G, bits, epoch_ctrl, stat);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

// Marks of lists of at least kPartBucketMinN items and kPartBucketMinK removals
// are grouped by bitmap region first: one random device atomic per removal
// runs at about 40 G/s, where the two bucketing kernels stream.  Whole calls
// on 4-byte items, random atomics -> bucketed: n=2^22, k=n/2 105 -> 43 us
// (clustered 160 -> 47 us); n=6,000,000, k=0.8n 202 -> 55 us; n=2^21,
// k=n/4 40 -> 36 us.  At n=2^20, k=2^18 the atomics still win (30 vs 41 us).
constexpr int64_t kPartBucketMinK = 1 << 19, kPartBucketMinN = 1 << 20;

// 1024 threads x 4 vectors: 64 KiB tiles (staged in LDS); bigger tiles take
// fewer look-back hops per byte (a 1 GiB probe: 16 KiB tiles 0.70 ms,
// 64 KiB 0.54 ms, 128 KiB 0.56 ms; 1024 x 4 2% ahead of 512 x 8).
constexpr int kVecWG = 1024, kVecVPT = 4;
template <typename T>
struct VecShape {
  static constexpr int VPT = kVecVPT;  // vectors per thread
  static constexpr int TILE = kVecWG * VPT * (16 / (int)sizeof(T));
};

template <typename T>
static rz_status partition_impl(rz_ctx* ctx, T* A, int64_t n, const int32_t* R, int64_t k, uint32_t flags,
                                hipStream_t s) {
  const bool vec = ((uintptr_t)A % 16) == 0;
  const int TILE = vec ? VecShape<T>::TILE : kPartWG * kPartIPT;
  const int64_t tiles = (n + TILE - 1) / TILE;
  rz_status st;
  if ((st = ensure(ctx, ctx->list_bits, (size_t)((n + 63) / 64) * 8, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->status, (size_t)tiles * 8, s)) != RZ_OK) return st;
  const int64_t groups = (tiles + 63) / 64;
  if (vec) {
    if ((st = ensure(ctx, ctx->ploaded, (size_t)tiles * 4, s)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->pgcnt, (size_t)groups * 4, s)) != RZ_OK) return st;
    if ((st = ensure(ctx, ctx->pgdone, (size_t)groups * 4, s)) != RZ_OK) return st;
  }
  uint32_t* bits = (uint32_t*)ctx->list_bits.ptr;
  uint32_t* ctrl = (uint32_t*)ctx->ctrl.ptr;
  const int64_t gw = (k + 255) / 256;
  const dim3 g1((uint32_t)(gw < (int64_t)ctx->cu_count * 8 ? (gw > 0 ? gw : 1) : ctx->cu_count * 8));
  if ((st = prof_mark(ctx, flags, 0, s)) != RZ_OK) return st;
  const bool bucketed = k >= kPartBucketMinK && n >= kPartBucketMinN && bucket_marks_fit(n);
  if (bucketed) {
    // the bitmap is zero here: every call leaves the words it used clear
    if ((st = bucket_marks(ctx, R, k, 0u, n, bits, ctrl, nullptr, s)) != RZ_OK) return st;
  } else {
    hipLaunchKernelGGL(rz_mark_all, g1, dim3(256), 0, s, R, (uint32_t)k, (uint32_t)n, bits, ctrl);
  }
  RZ_HIP(hipGetLastError());
  if ((st = prof_mark(ctx, flags, 1, s)) != RZ_OK) return st;
  if (vec)
    hipLaunchKernelGGL((rz_compact_vec<T, kVecWG, VecShape<T>::VPT>), dim3((uint32_t)tiles), dim3(kVecWG), 0, s, A,
                       (uint32_t)n, bits, (uint64_t*)ctx->status.ptr, (uint32_t*)ctx->ploaded.ptr,
                       (uint32_t*)ctx->pgcnt.ptr, (uint32_t*)ctx->pgdone.ptr, (uint32_t)tiles, ctrl);
  else
    hipLaunchKernelGGL((rz_compact<T, kPartWG, kPartIPT>), dim3((uint32_t)tiles), dim3(kPartWG), 0, s, A,
                       (uint32_t)n, bits, (uint64_t*)ctx->status.ptr, ctrl);
  RZ_HIP(hipGetLastError());
  if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
  return prof_mark(ctx, flags, 3, s);
}

rz_status launch_partition(rz_ctx* ctx, void* data, int64_t n, uint32_t esz, const int32_t* R, int64_t k,
                           uint32_t flags, hipStream_t s) {
  switch (esz) {
    case 4: return partition_impl<uint32_t>(ctx, (uint32_t*)data, n, R, k, flags, s);
    case 8: return partition_impl<uint64_t>(ctx, (uint64_t*)data, n, R, k, flags, s);
    case 16: return partition_impl<uint4>(ctx, (uint4*)data, n, R, k, flags, s);
    default: return RZ_EINVAL;
  }
}

}  // namespace rz
