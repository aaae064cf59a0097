// rz_batch.hip -- batched Redzone: many lists in one persistent launch that
// pulls work items from a device queue; also the single-launch form of the
// plain path (one list, items computed from the ticket).
//
// Work items of a list with k removals cut into m slices of L pairs:
//   SINGLE        (m == 1) one workgroup runs mark, move and pair for the
//                 whole list (the one-block algorithm, RemoveFromList2,
//                 redzone.hxx:236-444, on a 64-lane wavefront workgroup);
//   MARK  j       marks of removals [jL, (j+1)L) into the list's tail bitmap;
//   MOVE  j       slice j, after every MARK of the list has finished;
//   PAIR  j       share j of the orphan pairs, after every MOVE has finished;
//   FMOVE j       (fused items, batch variant 4, in place of MOVE and PAIR)
//                 slice j after every MARK of the list has finished: moves its
//                 pairs, keeps its orphans in LDS and pairs them through the
//                 list's look-back words and exchange words (move_pair_core,
//                 rz_fused_dev.hpp).  The list's MARK j also zeroes the
//                 slice's look-back word and its share of the list's exchange
//                 words, so no list state survives from an earlier launch.
// With batch variant 5, a list of 2^19 to 2^22 removals marks its tail in
// buckets instead of one atomic per tail removal:
//   MARK j        groups the tail removals of removals [jL, (j+1)L) by bucket
//                 (2^15 tail positions) in LDS, reserves room in each
//                 bucket's entry list with one atomic per bucket, and writes
//                 the entries out bucket by bucket (whole runs, not one
//                 scattered request per removal);
//   BUCKET b      after every MARK of the list, sets bucket b's bits in an
//                 LDS bitmap and writes its bitmap words whole; MOVE items wait
//                 for the BUCKET items.
// A list whose descriptor asks for the partition (method 3) runs the O(n)
// stable compaction instead:
//   PMARK j       marks removals [jL, (j+1)L) into the list's n-bit bitmap and
//                 zeroes a 1/m share of the list's tile status words;
//   PTILE j       tile j of the list's items, after every PMARK of the list
//                 has finished (compact_tile, rz_compact_dev.hpp): tiles are
//                 dequeued in order, so the in-place look-back pass never
//                 overwrites an item before its tile has read it.
// A batch is planned into one queue per XCD: lists go to XCDs by longest-
// processing-time order on their removal counts, and all items of a list sit
// in its XCD's queue.  Workgroup w drains queue w mod 8 first (the XCD the
// hardware dispatches it to), then helps with the others.  The holes of a list
// are then written from one XCD, which works on one or two lists at a time.
// On MI355X that made the 4096-list batch 1.7x faster than one shared queue,
// although the L2s still write back one 32-byte sector per hole: the saving is
// on the memory side, where the writes now arrive local in time and address.
// Within a queue the lists follow each other (MARK items of a list, then its
// MOVE items, then its PAIR items, with the next list's MARK items pulled
// ahead of the MOVE items and the PAIR items behind the next list's MOVE
// items; one-slice lists last), so each XCD writes into few lists at a time.  Items of a queue are dequeued in index order with
// an atomic ticket, and every wait is on items of the same list, which sit
// earlier in the same queue, so the queues drain whatever the number of
// resident workgroups: no cooperative launch is needed.
// The last workgroup to leave resets the tickets and the per-list counters, so
// the next launch (or graph replay) starts clean without a memset.
#include <string.h>

#include <algorithm>
#include <vector>

#include "rz_compact_dev.hpp"
#include "rz_fused_dev.hpp"
#include "rz_redzone_dev.hpp"

namespace rz {

constexpr int kQueueMaxSlices = 1024;  // slices per list in the queue kernel (LDS prefix)
enum ItemKind : uint32_t {
  kSingle = 0, kMark = 1, kMoveItem = 2, kPair = 3, kFusedMove = 4, kPMark = 5, kPTile = 6, kBucket = 7
};
// the item families a queue-kernel instantiation carries besides MARK, MOVE, PAIR and SINGLE
enum QueueMode : int { kModeFused = 1, kModePart = 2, kModeBuck = 4 };
// bucketed tail marks (batch variant 5)
constexpr uint32_t kBucketLog = 15, kBucketW = 1u << kBucketLog;  // tail positions per bucket (4 KB of bitmap)
constexpr uint32_t kMaxBuckets = 128;                             // lists of fewer than 2^22 removals
constexpr uint32_t kBucketMinK = 1u << 19;                        // ... and at least this many
constexpr uint32_t kBucketMaxL = 8192;                            // slice length the LDS staging holds
constexpr uint16_t kPlanBucketed = 5;                             // ListPlan::method of such a list
// longest slice of a list with fused items (its orphans take 4 bytes per pair of LDS)
constexpr uint32_t kFusedItemMaxL = 8192;

struct ListPlan {
  uint64_t data;      // T*
  uint64_t R;         // int32_t* removal list
  uint64_t Rd;        // int32_t* dest-orphan storage (R, or preserve scratch)
  uint32_t n, k, sZ, L;  // partition lists: L removals per PMARK item
  uint32_t m;            // slices (partition lists: PMARK items)
  uint16_t esz, method;  // method 3: the list runs PMARK and PTILE items
  uint32_t bits_off;     // first bitmap word of this list (k bits; partition lists n bits)
  uint32_t counts_off;   // first slice-count word of this list (partition lists: tile status words)
  uint32_t npair;        // number of PAIR items (0: the list runs fused items; partition lists: PTILE items)
  uint32_t slots_off;   // fused items: first exchange word of this list (k/2 + 1 words)
  uint64_t ent;          // bucketed marks: uint32_t* entry lists, kBucketW words per bucket
  uint64_t cur;          // bucketed marks: uint32_t* entry counts, one per bucket (zero between launches)
};
static_assert(sizeof(ListPlan) == 80, "ListPlan is one 80-byte record");

// queue control words (uint32 index into the queue control block): the ticket
// of queue x at kQTicket + 32 x (a line each), then the exit count
constexpr int kQueues = 8;
enum QueueWord : int { kQTicket = 0, kQExit = 32 * kQueues, kQWords = 32 * kQueues + 32 };
// queue x holds items [off[x], off[x + 1])
struct QueueSplit {
  uint32_t off[kQueues + 1];
};

// Bucketed marks of removals [a, b) of a list (MARK item of batch variant 5):
// the tail removals' positions z, grouped by bucket z >> kBucketLog in LDS,
// appended to their buckets' entry lists in whole runs.
template <int WG>
__device__ __forceinline__ void bucket_marks(const int32_t* __restrict__ R, uint32_t a, uint32_t b, uint32_t k,
                                             uint32_t sZ, uint32_t* __restrict__ ent, uint32_t* __restrict__ cur) {
  __shared__ uint32_t s_stage[kBucketMaxL];
  __shared__ uint32_t s_cnt[kMaxBuckets], s_lbase[kMaxBuckets], s_gbase[kMaxBuckets];
  __shared__ uint32_t s_total;
  const uint32_t tid = threadIdx.x;
  const uint32_t nb = (k + kBucketW - 1) >> kBucketLog;
  for (uint32_t w = tid; w < nb; w += WG) s_cnt[w] = 0u;
  __syncthreads();
  // U loads per thread in flight, then their uses (a use between them would
  // wait for each load); z = R[i] - sZ wraps to >= 2^31 > k below the tail,
  // and a slot past b loads sZ - 1, which maps to 0xffffffff
  constexpr int U = 8;
  for (uint32_t i0 = a + tid; i0 < b; i0 += WG * U) {
    uint32_t z[U];
#pragma unroll
    for (int u = 0; u < U; ++u) z[u] = i0 + u * WG < b ? (uint32_t)R[i0 + u * WG] : sZ - 1u;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t zz = z[u] - sZ;
      if (zz < k) atomicAdd(&s_cnt[zz >> kBucketLog], 1u);
    }
  }
  __syncthreads();
  if (tid < kWave) {  // exclusive prefix over the buckets, two per lane
    const uint32_t b0 = 2 * tid, b1 = b0 + 1;
    const uint32_t c0 = b0 < nb ? s_cnt[b0] : 0u, c1 = b1 < nb ? s_cnt[b1] : 0u;
    const uint32_t incl = wave_incl_scan(c0 + c1);
    const uint32_t ex = incl - c0 - c1;
    if (b0 < nb) s_lbase[b0] = ex;
    if (b1 < nb) s_lbase[b1] = ex + c0;
    if (tid == kWave - 1) s_total = incl;
  }
  __syncthreads();
  for (uint32_t w = tid; w < nb; w += WG) {  // room in each bucket's list: one atomic per bucket
    const uint32_t c = s_cnt[w];
    s_gbase[w] = c ? add_relaxed(&cur[w], c) : 0u;
    s_cnt[w] = 0u;
  }
  __syncthreads();
  for (uint32_t i0 = a + tid; i0 < b; i0 += WG * U) {  // the entries in LDS, bucket after bucket
    uint32_t z[U];
#pragma unroll
    for (int u = 0; u < U; ++u) z[u] = i0 + u * WG < b ? (uint32_t)R[i0 + u * WG] : sZ - 1u;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t zz = z[u] - sZ;
      if (zz < k) {
        const uint32_t bz = zz >> kBucketLog;
        s_stage[s_lbase[bz] + atomicAdd(&s_cnt[bz], 1u)] = zz;
      }
    }
  }
  __syncthreads();
  const uint32_t total = s_total;
  for (uint32_t t = tid; t < total; t += WG) {  // consecutive threads write consecutive entries of a bucket
    const uint32_t z = s_stage[t];
    const uint32_t bz = z >> kBucketLog;
    const uint32_t g = s_gbase[bz] + (t - s_lbase[bz]);
    if (g < kBucketW) st_relaxed(&ent[(size_t)bz * kBucketW + g], z);  // more only with duplicate removals
  }
}

// BUCKET item b: bucket b's bits into LDS, then its bitmap words written whole
// (the bitmap is zero outside the marks); the bucket's count back to zero.
template <int WG>
__device__ __forceinline__ void bucket_bits(uint32_t b, uint32_t k, const uint32_t* __restrict__ ent,
                                            uint32_t* __restrict__ cur, uint32_t* __restrict__ bits) {
  constexpr uint32_t WORDS = kBucketW / 32;
  __shared__ uint32_t s_bm[WORDS];
  const uint32_t tid = threadIdx.x;
  for (uint32_t w = tid; w < WORDS; w += WG) s_bm[w] = 0u;
  __syncthreads();
  const uint32_t c = min(ld_relaxed(&cur[b]), kBucketW);
  const uint32_t* e = ent + (size_t)b * kBucketW;
  const uint32_t lo = b * kBucketW;
  constexpr int U = 8;  // loads in flight per thread, then their uses
  for (uint32_t t0 = tid; t0 < c; t0 += WG * U) {
    uint32_t v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = t0 + u * WG < c ? ld_relaxed(&e[t0 + u * WG]) : lo - 1u;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t off = v[u] - lo;  // lo - 1 maps to 0xffffffff
      if (off < kBucketW) atomicOr(&s_bm[off >> 5], 1u << (off & 31));
    }
  }
  __syncthreads();
  const uint32_t w0 = lo >> 5, wend = (k + 31) >> 5;
  for (uint32_t w = tid; w < WORDS; w += WG) {
    const uint32_t v = s_bm[w];
    if (v && w0 + w < wend) st_relaxed(&bits[w0 + w], v);
  }
  if (tid == 0) st_relaxed(&cur[b], 0u);
}

// MODE (kModeFused, kModePart, kModeBuck): the item families this kernel
// instantiation runs besides MARK, MOVE, PAIR and SINGLE; the others leave that
// code out, which keeps their registers, LDS and occupancy.
template <typename T, int WG, int IPT, int MODE>
__device__ __forceinline__ void run_item(const ListPlan& P, uint32_t li, uint32_t kind, uint32_t j,
                                         uint32_t* __restrict__ done, uint32_t* __restrict__ bits,
                                         uint64_t* __restrict__ counts, uint64_t* __restrict__ slots,
                                         uint32_t* __restrict__ ctrl, MoveShared<WG, IPT>& msh,
                                         PairShared<WG, kQueueMaxSlices>& psh) {
  constexpr bool FUSED = (MODE & kModeFused) != 0, PART = (MODE & kModePart) != 0, BUCK = (MODE & kModeBuck) != 0;
  T* A = (T*)P.data;
  const int32_t* R = (const int32_t*)P.R;
  int32_t* Rd = (int32_t*)P.Rd;
  uint32_t* lb = bits + P.bits_off;
  uint64_t* lc = counts + P.counts_off;
  const uint32_t tid = threadIdx.x;
  if (kind == kSingle) {
    mark_body<false>(R, P.k, P.k, P.sZ, lb, nullptr, tid, WG);
    wait_all_memory();
    __syncthreads();
    move_slice<T, WG, IPT, true>(A, R, Rd, P.k, P.sZ, P.L, 0, lb, lc, msh);
    wait_all_memory();
    __syncthreads();
    pair_body<T, WG, kQueueMaxSlices, true>(A, Rd, P.k, P.sZ, P.L, 1, lc, ctrl, 0, 0, 1, psh);
  } else if (kind == kMark) {
    const uint32_t a = j * P.L, b = min(a + P.L, P.k);
    if (BUCK && P.method == kPlanBucketed) {
      if constexpr (BUCK) bucket_marks<WG>(R, a, b, P.k, P.sZ, (uint32_t*)P.ent, (uint32_t*)P.cur);
    } else {
      mark_body<false>(R + a, b - a, P.k, P.sZ, lb, nullptr, tid, WG);
    }
    if (FUSED && P.npair == 0) {
      // fused items: this slice's look-back word and a 1/m share of the
      // list's exchange words start at zero
      if (tid == 0) st_relaxed(&lc[j], 0ull);
      const uint32_t nslot = P.k / 2 + 1, share = (nslot + P.m - 1) / P.m;
      uint64_t* ls = slots + P.slots_off;
      const uint32_t lo = j * share, hi = min(lo + share, nslot);
      for (uint32_t w = lo + tid; w < hi; w += WG) st_relaxed(&ls[w], 0ull);
    }
    signal_done(&done[3 * li]);
  } else if (kind == kPMark) {
    if constexpr (PART) {
      const uint32_t a = j * P.L, b = min(a + P.L, P.k);
      for (uint32_t i = a + tid; i < b; i += WG) {
        const uint32_t r = (uint32_t)R[i];
        if (r < P.n) or_relaxed(&lb[r >> 5], 1u << (r & 31));
      }
      const uint32_t share = (P.npair + P.m - 1) / P.m;
      const uint32_t lo = j * share, hi = min(lo + share, P.npair);
      for (uint32_t w = lo + tid; w < hi; w += WG) st_relaxed(&lc[w], 0ull);
      signal_done(&done[3 * li]);
    }
  } else if (kind == kPTile) {
    if constexpr (PART) {
      wait_count(&done[3 * li], P.m, ctrl);
      compact_tile<T, WG, IPT, true>(A, P.n, lb, lc, j, 0u, ctrl);
    }
  } else if (kind == kFusedMove) {
    if constexpr (FUSED) {
      wait_count(&done[3 * li], P.m, ctrl);
      move_pair_core<T, WG, IPT, true, false>(A, R, P.k, P.sZ, P.L, P.m, lb, lc, slots + P.slots_off,
                                              P.k / 2 + 1, ctrl, 0, 0, j, 0u, 1u, nullptr);
    }
  } else if (kind == kBucket) {
    if constexpr (BUCK) {
      wait_count(&done[3 * li], P.m, ctrl);
      bucket_bits<WG>(j, P.k, (const uint32_t*)P.ent, (uint32_t*)P.cur, lb);
      signal_done(&done[3 * li + 2]);
    }
  } else if (kind == kMoveItem) {
    if (BUCK && P.method == kPlanBucketed)
      wait_count(&done[3 * li + 2], (P.k + kBucketW - 1) >> kBucketLog, ctrl);  // the BUCKET items
    else
      wait_count(&done[3 * li], P.m, ctrl);
    move_slice<T, WG, IPT, true>(A, R, Rd, P.k, P.sZ, P.L, j, lb, lc, msh, &done[3 * li + 1]);
  } else {
    wait_count(&done[3 * li + 1], P.m, ctrl);
    pair_body<T, WG, kQueueMaxSlices, true>(A, Rd, P.k, P.sZ, P.L, P.m, lc, ctrl, 0, j, P.npair, psh);
  }
}

// items == nullptr: one list (`single`), item q is MARK q (q < m), MOVE q-m
// (q < 2m) or PAIR q-2m; with m == 1 the only item is SINGLE.
template <int WG, int IPT, int MODE>
__global__ __launch_bounds__(WG) void rz_queue(ListPlan single, const ListPlan* __restrict__ plans,
                                                const uint2* __restrict__ items, QueueSplit qs, uint32_t chunk,
                                                uint32_t* __restrict__ qctl, uint32_t* __restrict__ done,
                                                uint32_t nlists, uint32_t* __restrict__ bits,
                                                uint64_t* __restrict__ counts, uint64_t* __restrict__ slots,
                                                uint32_t* __restrict__ ctrl) {
  __shared__ MoveShared<WG, IPT> msh;
  __shared__ PairShared<WG, kQueueMaxSlices> psh;
  __shared__ uint32_t s_q, s_last;
  const uint32_t tid = threadIdx.x;
  const uint32_t home = blockIdx.x & (kQueues - 1);
  for (uint32_t r = 0; r < (uint32_t)kQueues; ++r) {
    const uint32_t x = (home + r) & (kQueues - 1);
    uint32_t lo = 0, hi = 0;  // selects, not an indexed copy of the argument
#pragma unroll
    for (int i = 0; i < kQueues; ++i) {
      if ((uint32_t)i == x) lo = qs.off[i];
      if ((uint32_t)i == x) hi = qs.off[i + 1];
    }
    if (lo == hi) continue;
    for (;;) {
      if (tid == 0) s_q = lo + add_relaxed(&qctl[kQTicket + 32 * x], chunk);
      __syncthreads();
      const uint32_t q0 = s_q;
      __syncthreads();
      if (q0 >= hi) break;
      const uint32_t q1 = min(q0 + chunk, hi);
      for (uint32_t q = q0; q < q1; ++q) {
        uint32_t li, kind, j;
        ListPlan P;
        if (items) {
          const uint2 it = items[q];
          li = it.x & 0x0fffffffu;
          kind = it.x >> 28;
          j = it.y;
          P = plans[li];
        } else {
          P = single;
          li = 0;
          if (P.m == 1) {
            kind = kSingle;
            j = 0;
          } else if (q < P.m) {
            kind = kMark;
            j = q;
          } else if (q < 2 * P.m) {
            kind = kMoveItem;
            j = q - P.m;
          } else {
            kind = kPair;
            j = q - 2 * P.m;
          }
        }
        if (P.esz == 4)
          run_item<uint32_t, WG, IPT, MODE>(P, li, kind, j, done, bits, counts, slots, ctrl, msh, psh);
        else if (P.esz == 8)
          run_item<uint64_t, WG, IPT, MODE>(P, li, kind, j, done, bits, counts, slots, ctrl, msh, psh);
        else
          run_item<uint4, WG, IPT, MODE>(P, li, kind, j, done, bits, counts, slots, ctrl, msh, psh);
      }
    }
  }
  // the last workgroup out resets the queue state for the next launch
  if (tid == 0) {
    wait_all_memory();
    s_last = add_relaxed(&qctl[kQExit], 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (s_last) {
    for (uint32_t i = tid; i < 3 * nlists; i += WG) st_relaxed(&done[i], 0u);
    if (tid == 0) {
      for (int x = 0; x < kQueues; ++x) st_relaxed(&qctl[kQTicket + 32 * x], 0u);
      st_relaxed(&qctl[kQExit], 0u);
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
typedef void (*QueueKernel)(ListPlan, const ListPlan*, const uint2*, QueueSplit, uint32_t, uint32_t*, uint32_t*,
                            uint32_t, uint32_t*, uint64_t*, uint64_t*, uint32_t*);

// every item in queue 0 (one list, or a plan that is not split)
static QueueSplit one_queue(uint32_t nitems) {
  QueueSplit q;
  q.off[0] = 0;
  for (int x = 1; x <= kQueues; ++x) q.off[x] = nitems;
  return q;
}

template <int WG, int IPT>
static QueueKernel queue_kernel_of(int mode) {
  switch (mode & 7) {
    case 0: return rz_queue<WG, IPT, 0>;
    case 1: return rz_queue<WG, IPT, 1>;
    case 2: return rz_queue<WG, IPT, 2>;
    case 3: return rz_queue<WG, IPT, 3>;
    case 4: return rz_queue<WG, IPT, 4>;
    case 5: return rz_queue<WG, IPT, 5>;
    case 6: return rz_queue<WG, IPT, 6>;
    default: return rz_queue<WG, IPT, 7>;
  }
}
static QueueKernel queue_kernel(int wg, int ipt, int mode = 0) {
  if (wg == 256 && ipt == 4) return queue_kernel_of<256, 4>(mode);
  if (wg == 256 && ipt == 8) return queue_kernel_of<256, 8>(mode);
  if (wg == 512 && ipt == 4) return queue_kernel_of<512, 4>(mode);
  if (wg == 512 && ipt == 8) return queue_kernel_of<512, 8>(mode);
  return nullptr;
}

static rz_status queue_grid(rz_ctx* ctx, QueueKernel kern, int wg, size_t lds, uint32_t nitems, uint32_t* grid) {
  int per_cu = 0;
  RZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, wg, lds));
  if (per_cu < 1) per_cu = 1;
  uint32_t g = (uint32_t)per_cu * (uint32_t)ctx->cu_count;
  if (g > nitems) g = nitems;
  *grid = g < 1 ? 1 : g;
  return RZ_OK;
}

// Slice length of one list in the queue kernel: a multiple of one step, at
// least cfg.max_slices-derived length, and at most kQueueMaxSlices slices.
static void queue_slices(int64_t k, const rz_tile_config& cfg, uint32_t* L, uint32_t* m) {
  const int64_t step = (int64_t)cfg.wg_size * cfg.items;
  const int64_t steps = cfg.slice_steps > 0 ? cfg.slice_steps : 4;
  int64_t len = steps * step;
  const int64_t min_len = (k + kQueueMaxSlices - 1) / kQueueMaxSlices;
  if (len < min_len) len = (min_len + step - 1) / step * step;
  *L = (uint32_t)len;
  *m = (uint32_t)((k + len - 1) / len);
}

static void fill_plan(ListPlan& P, void* data, int32_t* R, int32_t* Rd, int64_t n, int64_t k, uint32_t esz,
                      const rz_tile_config& cfg) {
  memset(&P, 0, sizeof(P));
  P.data = (uint64_t)(uintptr_t)data;
  P.R = (uint64_t)(uintptr_t)R;
  P.Rd = (uint64_t)(uintptr_t)Rd;
  P.n = (uint32_t)n;
  P.k = (uint32_t)k;
  P.sZ = (uint32_t)(n - k);
  P.esz = esz;
  queue_slices(k, cfg, &P.L, &P.m);
  P.npair = P.m == 1 ? 1 : std::max<uint32_t>(1, (P.m + 3) / 4);
}

// Single list through the queue kernel (plain-path variant 2).
rz_status launch_queue_single(rz_ctx* ctx, void* data, int64_t n, uint32_t esz, int32_t* R, int64_t k,
                              uint32_t flags, const rz_tile_config& cfg, hipStream_t s) {
  QueueKernel kern = queue_kernel(cfg.wg_size, cfg.items);
  if (!kern) return RZ_EINVAL;
  rz_status st;
  int32_t* Rd = R;
  if (flags & RZ_PRESERVE_REMOVALS) {
    if ((st = ensure(ctx, ctx->orphans, (size_t)k * 4, s)) != RZ_OK) return st;
    Rd = (int32_t*)ctx->orphans.ptr;
  }
  ListPlan P;
  fill_plan(P, data, R, Rd, n, k, esz, cfg);
  if ((st = ensure(ctx, ctx->tail_bits, (size_t)((k + 63) / 64) * 8, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->counts, (size_t)P.m * 8, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->queue, 4096, s)) != RZ_OK) return st;
  const uint32_t nitems = P.m == 1 ? 1 : 2 * P.m + P.npair;
  uint32_t grid;
  if ((st = queue_grid(ctx, kern, cfg.wg_size, 0, nitems, &grid)) != RZ_OK) return st;
  uint32_t* qctl = (uint32_t*)ctx->queue.ptr;
  if ((st = prof_mark(ctx, flags, 0, s)) != RZ_OK) return st;
  if ((st = prof_mark(ctx, flags, 1, s)) != RZ_OK) return st;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(cfg.wg_size), 0, s, P, (const ListPlan*)nullptr, (const uint2*)nullptr,
                     one_queue(nitems), 1u, qctl, qctl + kQWords, 1u, (uint32_t*)ctx->tail_bits.ptr,
                     (uint64_t*)ctx->counts.ptr, (uint64_t*)nullptr, (uint32_t*)ctx->ctrl.ptr);
  RZ_HIP(hipGetLastError());
  if ((st = prof_mark(ctx, flags, 2, s)) != RZ_OK) return st;
  return prof_mark(ctx, flags, 3, s);
}

}  // namespace rz

// ---------------------------------------------------------------------------
// batch objects
// ---------------------------------------------------------------------------
struct rz_batch {
  rz_ctx* ctx = nullptr;
  int32_t nlists = 0;
  rz_tile_config cfg;
  uint32_t flags = 0;
  uint32_t nitems = 0, grid = 0, chunk = 1;
  rz::QueueKernel kern = nullptr;
  rz::QueueSplit split;
  rz::DeviceBuffer plans, qctl, bits, counts, orphans, slots, ents, curs;  // plans holds the items too
  const uint2* d_items = nullptr;  // the items, in `plans` after the list plans
  size_t lds = 0;  // dynamic LDS of the launch: the longest fused-item slice's orphan list
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  rz_batch_info info;
  uint32_t grid_cap = 0;  // resident workgroups of the queue kernel
  // host-side plan and pinned staging for asynchronous uploads
  std::vector<rz::ListPlan> h_plans;
  std::vector<uint2> h_items;
  std::vector<uint64_t> h_keys;
  std::vector<uint8_t> h_xcd;
  std::vector<uint32_t> h_seq;
  void* staging = nullptr;
  size_t staging_bytes = 0;
  hipEvent_t staged_ev = nullptr;
  bool staged = false;
};

using namespace rz;

static void free_buf(DeviceBuffer& b) {
  if (b.ptr) (void)hipFree(b.ptr);
  b.ptr = nullptr;
  b.bytes = 0;
}

// Plans the lists into batch b, reusing (growing) its device buffers.  The
// plan reaches the device synchronously, or (async) through pinned staging on
// stream s, ordered after the batch's previous launch on that stream.
static rz_status batch_plan(rz_batch* b, const rz_list_desc* lists, int32_t nlists, uint32_t flags,
                            const rz_tile_config& cfg, hipStream_t s, bool async) {
  rz_ctx* ctx = b->ctx;
  QueueKernel kern = queue_kernel(cfg.wg_size, cfg.items);
  if (!kern) return RZ_EINVAL;
  std::vector<ListPlan>& plans = b->h_plans;
  std::vector<uint2>& items = b->h_items;
  std::vector<uint64_t>& keys = b->h_keys;  // (k desc, index) for LPT ordering
  plans.resize((size_t)nlists);
  items.clear();
  keys.clear();
  uint64_t bits_words = 0, count_words = 0, orphan_words = 0, slot_words = 0, ent_words = 0, cur_words = 0;
  uint32_t fused_L = 0;
  bool any_part = false;
  for (int32_t i = 0; i < nlists; ++i) {
    const rz_list_desc& d = lists[i];
    if (d.n < 0 || d.k < 0 || d.k > d.n || d.n >= (1ll << 31) || esz_index(d.elem_bytes) < 0 ||
        (d.k > 0 && (!d.data || !d.removals)))
      return RZ_EINVAL;
    fill_plan(plans[i], d.data, d.removals, d.removals, d.n, d.k, d.elem_bytes, cfg);
    plans[i].bits_off = (uint32_t)bits_words;
    plans[i].counts_off = (uint32_t)count_words;
    if (flags & RZ_PRESERVE_REMOVALS) plans[i].Rd = orphan_words;  // offset, patched after allocation
    if (d.method == RZ_METHOD_PARTITION && d.k > 0 && d.k < d.n) {
      // the O(n) stable compaction: PMARK items of 4 tiles' worth of removals, one PTILE item per tile
      const uint32_t tile = (uint32_t)(cfg.wg_size * cfg.items);
      ListPlan& P = plans[i];
      P.method = RZ_METHOD_PARTITION;
      P.L = 4 * tile;
      P.m = (uint32_t)((d.k + P.L - 1) / P.L);
      P.npair = (uint32_t)((d.n + tile - 1) / tile);
      any_part = true;
      bits_words += (uint64_t)((d.n + 63) / 64) * 2;
      count_words += P.npair;
      keys.push_back(((uint64_t)(0xffffffffu - (uint32_t)d.k) << 32) | (uint32_t)i);
    } else if (d.k > 0 && d.k < d.n) {
      // batch variant 4: a list of several slices short enough for LDS runs fused items
      const uint64_t nslot = (uint64_t)d.k / 2 + 1;
      if (cfg.variant == 4 && plans[i].m > 1 && plans[i].L <= kFusedItemMaxL && slot_words + nslot < (1ull << 32)) {
        plans[i].npair = 0;
        plans[i].slots_off = (uint32_t)slot_words;
        slot_words += nslot;
        fused_L = std::max(fused_L, plans[i].L);
      } else if (cfg.variant == 5 && plans[i].m > 1 && d.k >= kBucketMinK && d.k < (1ll << 22) &&
                 plans[i].L <= kBucketMaxL) {
        // batch variant 5: bucketed tail marks (entry lists and counts, offsets patched after allocation)
        const uint32_t nb = (uint32_t)((d.k + kBucketW - 1) >> kBucketLog);
        plans[i].method = kPlanBucketed;
        plans[i].ent = ent_words;
        plans[i].cur = cur_words;
        ent_words += (uint64_t)nb * kBucketW;
        cur_words += nb;
      }
      bits_words += (uint64_t)((d.k + 63) / 64) * 2;
      count_words += plans[i].m;
      orphan_words += (uint64_t)d.k;
      keys.push_back(((uint64_t)(0xffffffffu - (uint32_t)d.k) << 32) | (uint32_t)i);
    } else {
      plans[i].m = 0;  // nothing to do
    }
  }
  if (bits_words >= (1ull << 32) || count_words >= (1ull << 32)) return RZ_EINVAL;
  std::sort(keys.begin(), keys.end());
  // lists to XCD queues, largest first onto the least loaded queue
  std::vector<uint8_t>& xcd = b->h_xcd;
  xcd.assign((size_t)nlists, 0);
  // queues in use: cfg.max_slices (1, 2 or 4) for batches, else one per XCD;
  // with fewer queues the device works on fewer lists at a time.  With
  // max_slices 16 every list of several slices has its items dealt over all 8
  // queues (item j of list l to queue (j + l) mod 8), in the same list order
  // in each, so that all XCDs work on the same list; the one-slice lists are
  // still assigned whole.  Each queue's order is then a subsequence of one
  // order of all items, in which an item's dependencies all come earlier, and
  // a workgroup leaves its home queue only once it is empty: the queues drain.
  const int nq = (cfg.max_slices == 1 || cfg.max_slices == 2 || cfg.max_slices == 4) ? cfg.max_slices : kQueues;
  const bool spread = cfg.max_slices == 16;
  uint64_t load[kQueues] = {};
  for (uint64_t key : keys) {
    const uint32_t li = (uint32_t)key;
    int best = 0;
    for (int x = 1; x < nq; ++x)
      if (load[x] < load[best]) best = x;
    xcd[li] = (uint8_t)best;
    // removals (a partition list: items / 16, its streaming cost on the same scale), plus a fixed cost per list
    load[best] += (plans[li].method == RZ_METHOD_PARTITION ? (uint64_t)plans[li].n / 16 : (uint64_t)plans[li].k) + 256u;
  }
  auto push = [&](uint32_t li, uint32_t kind, uint32_t j) { items.push_back(make_uint2(li | (kind << 28), j)); };
  QueueSplit split;
  for (int x = 0; x < kQueues; ++x) {
    split.off[x] = (uint32_t)items.size();
    // list after list, largest first: an XCD works on few lists at a time,
    // so its writes stay within one or two lists.
    // The next list's MARK items go ahead of this list's MOVE items, and a
    // list's PAIR items after the next list's MOVE items, so that a MOVE or
    // PAIR item rarely waits for the items it depends on.
    std::vector<uint32_t>& seq = b->h_seq;
    seq.clear();
    for (uint64_t key : keys) {
      const uint32_t li = (uint32_t)key;
      if ((spread || xcd[li] == x) && (plans[li].m > 1 || plans[li].method == RZ_METHOD_PARTITION)) seq.push_back(li);
    }
    auto push_all = [&](uint32_t li, uint32_t kind, uint32_t cnt) {
      for (uint32_t j = 0; j < cnt; ++j)
        if (!spread || (j + li) % kQueues == (uint32_t)x) push(li, kind, j);
    };
    // the three phases of a list: marks, then moves (or partition tiles), then PAIR items
    auto marks = [&](uint32_t li) {
      push_all(li, plans[li].method == RZ_METHOD_PARTITION ? kPMark : kMark, plans[li].m);
    };
    auto moves = [&](uint32_t li) {
      const ListPlan& P = plans[li];
      if (P.method == kPlanBucketed) push_all(li, kBucket, (P.k + kBucketW - 1) >> kBucketLog);
      if (P.method == RZ_METHOD_PARTITION) push_all(li, kPTile, P.npair);
      else push_all(li, P.npair ? kMoveItem : kFusedMove, P.m);
    };
    auto pairs = [&](uint32_t li) {
      if (plans[li].method != RZ_METHOD_PARTITION) push_all(li, kPair, plans[li].npair);
    };
    for (size_t i = 0; i < seq.size(); ++i) {
      if (i == 0) marks(seq[0]);
      if (i + 1 < seq.size()) marks(seq[i + 1]);
      moves(seq[i]);
      if (i >= 1) pairs(seq[i - 1]);
    }
    if (!seq.empty()) pairs(seq.back());
    for (uint64_t key : keys) {
      const uint32_t li = (uint32_t)key;
      if (xcd[li] == x && plans[li].m == 1 && plans[li].method != RZ_METHOD_PARTITION) push(li, kSingle, 0);
    }
  }
  split.off[kQueues] = (uint32_t)items.size();
  rz_status st;
  // the list plans and the items in one buffer, so that one copy uploads both
  const size_t plan_bytes = plans.size() * sizeof(ListPlan), item_bytes = items.size() * sizeof(uint2);
  const size_t item_off = (plan_bytes + 255) & ~(size_t)255, up_bytes = item_off + item_bytes;
  if ((st = ensure(ctx, b->plans, std::max<size_t>(64, up_bytes), s)) != RZ_OK ||
      (st = ensure(ctx, b->qctl, (size_t)(kQWords + 3 * (size_t)nlists + 64) * 4, s)) != RZ_OK ||
      (st = ensure(ctx, b->ents, std::max<size_t>(8, (size_t)ent_words * 4), s)) != RZ_OK ||
      (st = ensure(ctx, b->curs, std::max<size_t>(8, (size_t)cur_words * 4), s)) != RZ_OK ||
      (st = ensure(ctx, b->bits, std::max<size_t>(8, (size_t)bits_words * 4), s)) != RZ_OK ||
      (st = ensure(ctx, b->counts, std::max<size_t>(8, (size_t)count_words * 8), s)) != RZ_OK ||
      (st = ensure(ctx, b->slots, std::max<size_t>(8, (size_t)slot_words * 8), s)) != RZ_OK ||
      ((flags & RZ_PRESERVE_REMOVALS) &&
       (st = ensure(ctx, b->orphans, std::max<size_t>(8, (size_t)orphan_words * 4), s)) != RZ_OK))
    return st;
  if (flags & RZ_PRESERVE_REMOVALS)
    for (auto& P : plans) P.Rd = (uint64_t)(uintptr_t)((int32_t*)b->orphans.ptr + P.Rd);
  for (auto& P : plans)
    if (P.method == kPlanBucketed) {
      P.ent = (uint64_t)(uintptr_t)((uint32_t*)b->ents.ptr + P.ent);
      P.cur = (uint64_t)(uintptr_t)((uint32_t*)b->curs.ptr + P.cur);
    }
  b->d_items = (const uint2*)((char*)b->plans.ptr + item_off);
  if (async) {
    const size_t need = up_bytes;
    if (b->staged) RZ_HIP(hipEventSynchronize(b->staged_ev));  // the previous upload has left the staging area
    if (need > b->staging_bytes) {
      if (b->staging) RZ_HIP(hipHostFree(b->staging));
      b->staging = nullptr;
      size_t cap = 1 << 16;
      while (cap < need) cap *= 2;
      if (hipHostMalloc(&b->staging, cap, hipHostMallocDefault) != hipSuccess) return RZ_ENOMEM;
      b->staging_bytes = cap;
    }
    if (!b->staged_ev) RZ_HIP(hipEventCreateWithFlags(&b->staged_ev, hipEventDisableTiming));
    memcpy(b->staging, plans.data(), plan_bytes);
    memcpy((char*)b->staging + item_off, items.data(), item_bytes);
    if (up_bytes) RZ_HIP(hipMemcpyAsync(b->plans.ptr, b->staging, up_bytes, hipMemcpyHostToDevice, s));
    RZ_HIP(hipEventRecord(b->staged_ev, s));
    b->staged = true;
  } else {
    if (plan_bytes) RZ_HIP(hipMemcpy(b->plans.ptr, plans.data(), plan_bytes, hipMemcpyHostToDevice));
    if (item_bytes) RZ_HIP(hipMemcpy((char*)b->plans.ptr + item_off, items.data(), item_bytes, hipMemcpyHostToDevice));
  }
  const size_t lds = (size_t)fused_L * 4;
  // some list runs fused items, or partition items
  const int mode = (fused_L ? kModeFused : 0) | (any_part ? kModePart : 0) | (ent_words ? kModeBuck : 0);
  if (mode) kern = queue_kernel(cfg.wg_size, cfg.items, mode);
  if (kern != b->kern || lds != b->lds) {
    b->kern = kern;
    b->lds = lds;
    b->grid_cap = 0;
  }
  if (!b->grid_cap && (st = queue_grid(ctx, kern, cfg.wg_size, lds, 0xffffffffu, &b->grid_cap)) != RZ_OK) return st;
  b->nlists = nlists;
  b->cfg = cfg;
  b->flags = flags;
  b->nitems = (uint32_t)items.size();
  b->split = split;
  b->chunk = cfg.chunk > 0 ? (uint32_t)cfg.chunk : 1u;
  b->grid = std::max<uint32_t>(1, std::min<uint32_t>(b->grid_cap, b->nitems));
  b->info.nitems = b->nitems;
  b->info.grid = b->grid;
  b->info.bitmap_bytes = (int64_t)bits_words * 4;
  b->info.lists_active = (int32_t)keys.size();
  return RZ_OK;
}

// After a failed device check: the batch's bitmap, slice counts and queue
// words back to zero (the launch keeps them zero when every wait succeeds).
rz_status rz::batch_reset_scratch(rz_batch* b) {
  DeviceBuffer* bufs[] = {&b->qctl, &b->bits, &b->counts, &b->curs};
  for (DeviceBuffer* d : bufs)
    if (d->ptr) RZ_HIP(hipMemset(d->ptr, 0, d->bytes));
  return RZ_OK;
}

void rz::batch_detach(rz_batch* b) { b->ctx = nullptr; }

// Plans into the context's reusable batch and launches it (rz_remove_batched).
rz_status rz::batch_plan_and_launch(rz_ctx* ctx, const rz_list_desc* lists, int32_t nlists, uint32_t flags,
                                hipStream_t s) {
  if (!ctx->last_batch) {
    ctx->last_batch = new rz_batch();
    ctx->last_batch->ctx = ctx;
  }
  rz_status st = batch_plan(ctx->last_batch, lists, nlists, flags, ctx->batch_cfg, s, true);
  if (st != RZ_OK) return st;
  return rz_batch_launch(ctx->last_batch, s);
}

extern "C" {

rz_status rz_batch_create(rz_ctx* ctx, const rz_list_desc* lists, int32_t nlists, uint32_t flags,
                          const rz_tile_config* cfg_in, rz_batch** out) {
  if (!ctx || !out || nlists < 0 || (nlists > 0 && !lists)) return RZ_EINVAL;
  if (nlists >= (1 << 28)) return RZ_EINVAL;
  *out = nullptr;
  const rz_tile_config cfg = cfg_in ? *cfg_in : ctx->batch_cfg;
  if (!queue_kernel(cfg.wg_size, cfg.items)) return RZ_EINVAL;
  int dev_prev = -1;
  (void)hipGetDevice(&dev_prev);
  if (dev_prev != ctx->device) (void)hipSetDevice(ctx->device);
  rz_batch* b = new rz_batch();
  b->ctx = ctx;
  const rz_status st = batch_plan(b, lists, nlists, flags, cfg, nullptr, false);
  if (dev_prev >= 0 && dev_prev != ctx->device) (void)hipSetDevice(dev_prev);
  if (st != RZ_OK) {
    rz_batch_destroy(b);
    return st;
  }
  {
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->batches.push_back(b);
  }
  *out = b;
  return RZ_OK;
}

rz_status rz_batch_launch(rz_batch* b, void* stream) {
  if (!b || !b->ctx) return RZ_EINVAL;
  if (b->nitems == 0) return RZ_OK;
  hipStream_t s = (hipStream_t)stream;
  if (b->exec) {
    RZ_HIP(hipGraphLaunch(b->exec, s));
    return RZ_OK;
  }
  uint32_t* q = (uint32_t*)b->qctl.ptr;
  hipLaunchKernelGGL(b->kern, dim3(b->grid), dim3(b->cfg.wg_size), b->lds, s, ListPlan{},
                     (const ListPlan*)b->plans.ptr, b->d_items, b->split, b->chunk, q,
                     q + kQWords, (uint32_t)b->nlists, (uint32_t*)b->bits.ptr, (uint64_t*)b->counts.ptr,
                     (uint64_t*)b->slots.ptr, (uint32_t*)b->ctx->ctrl.ptr);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

rz_status rz_batch_capture(rz_batch* b, void* stream) {
  if (!b || !b->ctx) return RZ_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  if (b->graph) {
    (void)hipGraphDestroy(b->graph);
    b->graph = nullptr;
  }
  RZ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(b->kern, dim3(b->grid), dim3(b->cfg.wg_size), b->lds, s, ListPlan{},
                     (const ListPlan*)b->plans.ptr, b->d_items, b->split, b->chunk,
                     (uint32_t*)b->qctl.ptr, (uint32_t*)b->qctl.ptr + kQWords, (uint32_t)b->nlists,
                     (uint32_t*)b->bits.ptr, (uint64_t*)b->counts.ptr, (uint64_t*)b->slots.ptr,
                     (uint32_t*)b->ctx->ctrl.ptr);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(s, &g);
  if (e != hipSuccess || !g) return RZ_EHIP;
  b->graph = g;
  RZ_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  return RZ_OK;
}

rz_status rz_batch_get_info(rz_batch* b, rz_batch_info* out) {
  if (!b || !out) return RZ_EINVAL;
  *out = b->info;
  return RZ_OK;
}

rz_status rz_batch_destroy(rz_batch* b) {
  if (!b) return RZ_OK;
  if (b->ctx && b->ctx->last_batch != b) {
    std::lock_guard<std::mutex> lock(b->ctx->mu);
    std::vector<rz_batch*>& v = b->ctx->batches;
    v.erase(std::remove(v.begin(), v.end(), b), v.end());
  }
  (void)hipDeviceSynchronize();
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  free_buf(b->plans);
  free_buf(b->qctl);
  free_buf(b->bits);
  free_buf(b->counts);
  free_buf(b->orphans);
  free_buf(b->slots);
  free_buf(b->ents);
  free_buf(b->curs);
  if (b->staging) (void)hipHostFree(b->staging);
  if (b->staged_ev) (void)hipEventDestroy(b->staged_ev);
  delete b;
  return RZ_OK;
}

}  // extern "C"
