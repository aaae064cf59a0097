// rz_dense.hip -- region-bucketed Redzone for dense removal lists (variant 5).
//
// A scattered 4-byte write costs one read-modify-write of a DRAM burst, so the
// move step of the plain forms runs at about 25 G holes/s whatever the
// density, and so does the atomic tail marking.  When removals are dense
// (k/n of several percent and up) most bursts hold several targets, and
// writing them while the burst's line sits in one L2 turns many
// read-modify-writes into one.  This form regroups both the tail marks and the
// head writes by address before issuing them, in streaming passes:
//
//   tail marks (cfg.chunk == 2; otherwise the atomic rz_mark_tail)
//     bucket_marks (rz_partition.hip) over the window [sZ, n): the red-zone
//               removals appended to per-bin entry lists, then each bin's
//               bits set in an LDS bitmap and stored (two kernels)
//   head writes
//     count     tile j: per coarse destination bin, the number of head writes
//               (moves and dest orphans) and the tile's dest / source orphan
//               counts
//     colscan / topscan   coarse bin bases, per-tile offsets, per-tile orphan
//               ranks (exclusive, in i order)
//     entries   tile j again, in steps staged through LDS: (position, value)
//               runs of the moves per coarse bin; the t-th dest orphan's
//               position goes to dpos[t] and the t-th source orphan's value
//               to src[t].  Orphan ranks follow i, so the pairing is the
//               canonical one (redzone.hxx:834-934).
//     refine    chunks of the coarse runs, staged through LDS by fine bin
//               (about 1 MiB of the list each); runs go to fixed-capacity fine
//               bins through one atomic cursor add per (chunk, fine bin).
//               Distinct removals write each position at most once, so a fine
//               bin never holds more entries than it has positions.
//     refine    again, on the orphan pairs (dpos[t], src[t]) read in rank
//               order, into the same fine bins
//     scatter   each fine bin is split among G workgroups laid out to share
//               one XCD (workgroups are dispatched to the 8 XCDs round robin),
//               bins taken in order, so each XCD writes about two bins at a
//               time and the writes combine in its L2: A[d] = value.
//
// Counts and offsets are recomputed on every call.  The fine cursors are
// zeroed by the count kernel of the next call; the tail bitmap is zero on
// entry, set by the tail marks (or rz_mark_tail) and cleared by the entries
// kernel as it reads it.  With an invalid list the passes still
// classify identically, every index is checked against its buffer, and the
// orphan-count mismatch (or a fine-bin overflow) is flagged.
#include <atomic>

#include "rz_redzone_dev.hpp"

namespace rz {

constexpr int kCoarseMax = 1024;      // coarse destination bins and tail bins
constexpr int kFineMax = 4096;        // fine destination bins
constexpr uint32_t kTailBinLog = 19;  // the tail marks' largest bin: k up to kCoarseMax << kTailBinLog = 2^29
constexpr uint32_t kCoarseTarget = 64;

__device__ __forceinline__ uint32_t classify(uint32_t d, bool dead, uint32_t sZ) {
  const bool dst_ok = d < sZ;
  return dead ? (dst_ok ? kDstOrphan : kNone) : (dst_ok ? kMove : kSrcOrphan);
}

// Exclusive scan of a[0, N) in LDS, in place (N <= kFineMax); returns the total.
template <int WG>
__device__ uint32_t lds_excl_scan(uint32_t* a, uint32_t N, uint32_t* tmp) {
  constexpr uint32_t MAXPER = kFineMax / WG;
  const uint32_t per = (N + WG - 1) / WG;
  const uint32_t t = threadIdx.x;
  uint32_t loc[MAXPER], sum = 0;
#pragma unroll
  for (uint32_t q = 0; q < MAXPER; ++q) {
    const uint32_t idx = t * per + q;
    loc[q] = (q < per && idx < N) ? a[idx] : 0u;
    sum += loc[q];
  }
  uint32_t total;
  uint32_t ex = block_excl_scan<WG>(sum, tmp, &total);
#pragma unroll
  for (uint32_t q = 0; q < MAXPER; ++q) {
    const uint32_t idx = t * per + q;
    if (q < per && idx < N) {
      a[idx] = ex;
      ex += loc[q];
    }
  }
  __syncthreads();
  return total;
}


// ---- count ------------------------------------------------------------------
// Also zeroes the fine-bin cursors that this call's refine kernel adds to.
template <int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_dense_count(const int32_t* __restrict__ R, uint32_t k, uint32_t sZ,
                                                      uint32_t L, const uint32_t* __restrict__ bits, uint32_t NB,
                                                      uint32_t bs, uint32_t* __restrict__ cnt,
                                                      uint64_t* __restrict__ ocnt, uint32_t* __restrict__ fcur,
                                                      uint32_t NBf) {
  __shared__ uint32_t hist[kCoarseMax];
  __shared__ uint32_t s_nd, s_ns;
  const uint32_t tid = threadIdx.x, j = blockIdx.x;
  const uint32_t a = j * L, b = min(a + L, k);
  for (uint32_t q = j * WG + tid; q < NBf; q += gridDim.x * WG) fcur[q] = 0u;
  for (uint32_t q = tid; q < NB; q += WG) hist[q] = 0u;
  if (tid == 0) { s_nd = 0u; s_ns = 0u; }
  __syncthreads();
  uint32_t nd = 0, ns = 0;
  for (uint32_t base = a; base < b; base += WG * IPT) {
    uint32_t d[IPT], wv[IPT];
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t p = base + q * WG + tid;
      d[q] = sZ;
      wv[q] = 0xffffffffu;
      if (p < b) { d[q] = (uint32_t)R[p]; wv[q] = bits[p >> 5]; }
    }
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t p = base + q * WG + tid;
      const uint32_t cls = classify(d[q], (wv[q] >> (p & 31)) & 1u, sZ);
      lds_bin_count(hist, d[q] >> bs, cls == kMove);
      nd += cls == kDstOrphan;
      ns += cls == kSrcOrphan;
    }
  }
  nd = wave_sum(nd);
  ns = wave_sum(ns);
  if (lane_id() == 0) { atomicAdd(&s_nd, nd); atomicAdd(&s_ns, ns); }
  __syncthreads();
  for (uint32_t q = tid; q < NB; q += WG) cnt[(size_t)j * NB + q] = hist[q];
  if (tid == 0) ocnt[j] = (uint64_t)s_nd | ((uint64_t)s_ns << 32);
}

// ---- colscan: per bin, exclusive prefix over tiles --------------------------
// Workgroup g (1024 threads) owns bins [32g, 32g+32): thread t takes bin
// 32g + t%32 and the t/32-th chunk of tiles.  off[j*NB + bin] gets the prefix.
__global__ __launch_bounds__(1024) void rz_dense_colscan(const uint32_t* __restrict__ cnt, uint32_t nt, uint32_t NB,
                                                        uint32_t* __restrict__ off, uint32_t* __restrict__ btot) {
  __shared__ uint32_t part[32][33];
  const uint32_t t = threadIdx.x, lane = t & 31u, c = t >> 5;
  const uint32_t bin = blockIdx.x * 32u + lane;
  const uint32_t tpc = (nt + 31u) / 32u;
  const uint32_t j0 = c * tpc, j1 = min(j0 + tpc, nt);
  uint32_t s = 0;
  for (uint32_t jj = j0; jj < j1; ++jj) s += cnt[(size_t)jj * NB + bin];
  part[c][lane] = s;
  __syncthreads();
  if (c == 0) {
    uint32_t run = 0;
    for (int q = 0; q < 32; ++q) {
      const uint32_t v = part[q][lane];
      part[q][lane] = run;
      run += v;
    }
    btot[bin] = run;
  }
  __syncthreads();
  uint32_t run = part[c][lane];
  for (uint32_t jj = j0; jj < j1; ++jj) {
    const uint32_t v = cnt[(size_t)jj * NB + bin];
    off[(size_t)jj * NB + bin] = run;
    run += v;
  }
}

// ---- topscan: bin bases and per-tile orphan ranks (one workgroup) ----------
__global__ __launch_bounds__(1024) void rz_dense_topscan(const uint32_t* __restrict__ btot, uint32_t NB,
                                                        uint32_t* __restrict__ binbase,
                                                        const uint64_t* __restrict__ ocnt, uint32_t nt,
                                                        uint64_t* __restrict__ obase, uint32_t* __restrict__ ctrl,
                                                        int stats) {
  __shared__ uint32_t tmp32[1024 / kWave + 1];
  __shared__ uint64_t tmp64[1024 / kWave + 1];
  const uint32_t t = threadIdx.x;
  {
    const uint32_t v = t < NB ? btot[t] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan<1024>(v, tmp32, &total);
    if (t < NB) binbase[t] = ex;
    if (t == 0) binbase[NB] = total;
  }
  if (!obase) return;
  uint64_t carry = 0;
  for (uint32_t c0 = 0; c0 < nt; c0 += 1024u * 4u) {
    uint64_t loc[4], sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t jj = c0 + t * 4u + q;
      loc[q] = jj < nt ? ocnt[jj] : 0ull;
      sum += loc[q];
    }
    uint64_t chunk;
    uint64_t ex = carry + block_excl_scan<1024>(sum, tmp64, &chunk);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t jj = c0 + t * 4u + q;
      if (jj < nt) obase[jj] = ex;
      ex += loc[q];
    }
    carry += chunk;
  }
  if (t == 0) {
    obase[nt] = carry;
    const uint32_t td = (uint32_t)carry, ts = (uint32_t)(carry >> 32);
    obase[nt + 1] = min(td, ts);  // orphan pairs (the orphan refine's entry count, low word)
    if (td != ts) or_relaxed(&ctrl[kCtrlError], kErrOrphanMismatch);
    if (stats) st_relaxed((uint64_t*)&ctrl[kCtrlStatOrphans], (uint64_t)min(td, ts));
  }
}

// ---- entries: staged (position, value) runs per coarse bin ------------------
// Dynamic LDS: S values, S positions, S bin ids (S = WG * IPT).
template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_dense_entries(const T* __restrict__ A, const int32_t* __restrict__ R,
                                                        uint32_t k, uint32_t sZ, uint32_t L,
                                                        uint32_t* __restrict__ bits, uint32_t NB, uint32_t bs,
                                                        const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ binbase,
                                                        const uint64_t* __restrict__ obase,
                                                        uint32_t* __restrict__ ent_d, T* __restrict__ ent_v,
                                                        T* __restrict__ src, uint32_t* __restrict__ dpos,
                                                        uint32_t src_cap) {
  constexpr int NW = WG / kWave;
  constexpr int E = MoveShared<WG, IPT>::E;
  constexpr int PER = (E + kWave - 1) / kWave;
  constexpr uint32_t S = WG * IPT;
  extern __shared__ __align__(16) unsigned char dyn[];
  T* sv = (T*)dyn;
  uint32_t* sd = (uint32_t*)(sv + S);
  uint16_t* sb = (uint16_t*)(sd + S);
  __shared__ MoveShared<WG, IPT> sh;
  __shared__ uint32_t lcnt[kCoarseMax + 1], gb[kCoarseMax], gcur[kCoarseMax];
  __shared__ uint32_t tmp[WG / kWave + 1];
  uint32_t* s_off = sh.off;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint32_t j = blockIdx.x;
  const uint32_t a = j * L, b = min(a + L, k);
  const T* __restrict__ Z = A + sZ;
  for (uint32_t q = tid; q < NB; q += WG) {
    gcur[q] = binbase[q] + off[(size_t)j * NB + q];
    lcnt[q] = 0u;
  }
  const uint64_t ob = obase[j];
  const uint32_t xd = (uint32_t)ob, xs = (uint32_t)(ob >> 32);
  __syncthreads();
  uint32_t carry_s = 0, carry_d = 0;
  for (uint32_t base = a; base < b; base += S) {
    uint32_t d[IPT], wv[IPT], info[IPT], rk[IPT];
    T z[IPT];
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t p = base + q * WG + tid;
      d[q] = sZ;
      z[q] = T();  // defined in every lane: keeps 16-byte items in registers
      wv[q] = 0xffffffffu;
      if (p < b) {
        d[q] = (uint32_t)R[p];
        z[q] = Z[p];
        wv[q] = bits[p >> 5];
      }
    }
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t p = base + q * WG + tid;
      const uint32_t cls = classify(d[q], (wv[q] >> (p & 31)) & 1u, sZ);
      const uint64_t ms = ballot(cls == kSrcOrphan);
      const uint64_t md = ballot(cls == kDstOrphan);
      info[q] = cls | (lane_rank(ms) << 8) | (lane_rank(md) << 16);
      if (lane == 0) s_off[q * NW + wave] = (uint32_t)__popcll(ms) | ((uint32_t)__popcll(md) << 16);
      rk[q] = lds_bin_add(lcnt, d[q] >> bs, cls == kMove);
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
      if (lane == kWave - 1) s_off[E] = x;
    }
    // reserve this step's run of every bin
    for (uint32_t q = tid; q < NB; q += WG) {
      gb[q] = gcur[q];
      gcur[q] += lcnt[q];
    }
    __syncthreads();
    const uint32_t total = lds_excl_scan<WG>(lcnt, NB, tmp);
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t cls = info[q] & 0xffu;
      const uint32_t o = s_off[q * NW + wave];
      if (cls == kSrcOrphan) {
        const uint32_t t = xs + carry_s + (o & 0xffffu) + ((info[q] >> 8) & 0xffu);
        if (t < src_cap) src[t] = z[q];
      } else if (cls == kDstOrphan) {
        const uint32_t t = xd + carry_d + (o >> 16) + (info[q] >> 16);
        if (t < src_cap) dpos[t] = d[q];
      } else if (cls == kMove) {
        const uint32_t bin = d[q] >> bs;
        const uint32_t e = lcnt[bin] + rk[q];
        sd[e] = d[q];
        sv[e] = z[q];
        sb[e] = (uint16_t)bin;
      }
    }
    const uint32_t tot = s_off[E];
    carry_s += tot & 0xffffu;
    carry_d += tot >> 16;
    __syncthreads();
    for (uint32_t e = tid; e < total; e += WG) {
      const uint32_t bin = sb[e];
      const uint32_t g = gb[bin] + (e - lcnt[bin]);
      ent_d[g] = sd[e];
      ent_v[g] = sv[e];
    }
    __syncthreads();
    for (uint32_t q = tid; q < NB; q += WG) lcnt[q] = 0u;
    __syncthreads();
  }
  // leave the tail bitmap zero for the next call (tiles are word aligned)
  for (uint32_t w = (a >> 5) + tid; w < ((b + 31u) >> 5); w += WG) bits[w] = 0u;
}

// ---- refine: coarse runs -> fixed-capacity fine bins -------------------------
// Chunk c = entries [c*S, (c+1)*S) of the coarse arrays (the grid is sized for
// k entries; chunks past the total leave at once).  Fine bin f holds the
// entries of positions [f << bsf, (f+1) << bsf) at fent[(f << bsf) + ...].
template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_dense_refine(const uint32_t* __restrict__ ent_d,
                                                       const T* __restrict__ ent_v,
                                                       const uint32_t* __restrict__ total_p,
                                                       uint32_t* __restrict__ fent_d, T* __restrict__ fent_v,
                                                       uint32_t* __restrict__ fcur, uint32_t NBf, uint32_t bsf,
                                                       uint32_t* __restrict__ ctrl) {
  constexpr uint32_t S = WG * IPT;
  extern __shared__ __align__(16) unsigned char dyn[];
  T* sv = (T*)dyn;
  uint32_t* sd = (uint32_t*)(sv + S);
  uint16_t* sb = (uint16_t*)(sd + S);
  __shared__ uint32_t lcnt[kFineMax + 1], gb[kFineMax];
  __shared__ uint32_t tmp[WG / kWave + 1];
  const uint32_t tid = threadIdx.x;
  const uint32_t total = *total_p;
  const uint32_t c0 = blockIdx.x * S;
  if (c0 >= total) return;  // uniform across the workgroup
  const uint32_t c1 = min(c0 + S, total);
  for (uint32_t q = tid; q < NBf; q += WG) lcnt[q] = 0u;
  __syncthreads();
  uint32_t d[IPT], rk[IPT];
  T v[IPT];
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const uint32_t e = c0 + q * WG + tid;
    d[q] = 0xffffffffu;
    v[q] = T();
    if (e < c1) {
      d[q] = ent_d[e];
      v[q] = ent_v[e];
    }
  }
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    const uint32_t f = d[q] >> bsf;
    const bool valid = d[q] != 0xffffffffu && f < NBf;
    rk[q] = lds_bin_add(lcnt, f, valid);
    if (!valid) rk[q] = 0xffffffffu;
  }
  __syncthreads();
  for (uint32_t q = tid; q < NBf; q += WG) {
    const uint32_t cq = lcnt[q];
    gb[q] = cq ? add_relaxed(&fcur[q], cq) : 0u;
  }
  __syncthreads();
  const uint32_t tot = lds_excl_scan<WG>(lcnt, NBf, tmp);
#pragma unroll
  for (int q = 0; q < IPT; ++q) {
    if (rk[q] != 0xffffffffu) {
      const uint32_t f = d[q] >> bsf;
      const uint32_t e = lcnt[f] + rk[q];
      sd[e] = d[q];
      sv[e] = v[q];
      sb[e] = (uint16_t)f;
    }
  }
  __syncthreads();
  const uint32_t cap = 1u << bsf;
  bool over = false;
  for (uint32_t e = tid; e < tot; e += WG) {
    const uint32_t f = sb[e];
    const uint32_t r = gb[f] + (e - lcnt[f]);
    if (r < cap) {
      const size_t g = ((size_t)f << bsf) + r;
      fent_d[g] = sd[e];
      fent_v[g] = sv[e];
    } else {
      over = true;
    }
  }
  if (over) or_relaxed(&ctrl[kCtrlError], kErrOrphanMismatch);
}

// ---- scatter ----------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void rz_dense_scatter(T* __restrict__ A, const uint32_t* __restrict__ fent_d,
                                                        const T* __restrict__ fent_v,
                                                        const uint32_t* __restrict__ fcur, uint32_t NBf,
                                                        uint32_t bsf, uint32_t G, uint32_t sZ) {
  const uint32_t w = blockIdx.x;
  const uint32_t x = w & 7u, rest = w >> 3;
  const uint32_t g = rest % G, bin = (rest / G) * 8u + x;
  if (bin >= NBf) return;
  const uint32_t cnt = min(fcur[bin], 1u << bsf);
  const size_t s = (size_t)bin << bsf;
  const uint32_t per = (cnt + G - 1) / G;
  const uint32_t lo = g * per, hi = min(lo + per, cnt);
  for (uint32_t q = lo + threadIdx.x; q < hi; q += 1024u) {
    const uint32_t d = fent_d[s + q];
    if (d < sZ) A[d] = fent_v[s + q];
  }
}

// ---------------------------------------------------------------------------
// launch
// ---------------------------------------------------------------------------
struct DenseShape {
  uint32_t L, nt;     // tiles of the removal list
  uint32_t NB, bs;    // coarse destination bins (multiple of 32)
  uint32_t NBf, bsf;  // fine destination bins (multiple of 8)
  uint32_t NBt;       // tail bins (multiple of 32), 0 = atomic marks
  uint32_t G;         // scatter workgroups per fine bin
};

static uint32_t bin_log(int64_t span, int64_t target) {
  uint32_t lg = 10;
  while (((span + ((int64_t)1 << lg) - 1) >> lg) > target) ++lg;
  return lg;
}

// Dynamic-LDS limits are set with hipFuncSetAttribute once per device (the
// current one, which every entry point selects): bit d of the mask.
static bool attr_needed(const std::atomic<uint64_t>& mask, int dev) {
  return dev < 0 || dev >= 64 || !(mask.load(std::memory_order_relaxed) & (1ull << dev));
}
static void attr_done(std::atomic<uint64_t>& mask, int dev) {
  if (dev >= 0 && dev < 64) mask.fetch_or(1ull << dev, std::memory_order_relaxed);
}

// small buffer: [btot | binbase (+1) | tbtot | tbinbase (+1) | fcur] (uint32,
// kSmallStride each, fcur kFineMax + 64), then [ocnt | obase (+1)] (uint64)
constexpr size_t kSmallStride = kCoarseMax + 64;
constexpr size_t kSmallWords = 4 * kSmallStride + kFineMax + 64;
constexpr size_t kSmallBytes = kSmallWords * 4 + 2 * ((size_t)kMaxSlices + 64) * 8;

template <typename T, int WG, int IPT>
static rz_status dense_launch(rz_ctx* ctx, T* A, const int32_t* R, uint32_t k, uint32_t sZ, const DenseShape& d,
                              uint32_t* bits, uint32_t* ctrl, int stats, uint32_t src_cap, hipStream_t s) {
  constexpr uint32_t S = WG * IPT;
  const size_t lds_ent = (size_t)S * (sizeof(T) + 4 + 2);
  static std::atomic<uint64_t> attr_mask{0};
  if (attr_needed(attr_mask, ctx->device)) {
    RZ_HIP(hipFuncSetAttribute((const void*)rz_dense_entries<T, WG, IPT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_ent));
    RZ_HIP(hipFuncSetAttribute((const void*)rz_dense_refine<T, WG, IPT>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)lds_ent));
    attr_done(attr_mask, ctx->device);
  }
  uint32_t* cnt = (uint32_t*)ctx->dense_cnt.ptr;
  uint32_t* btot = (uint32_t*)ctx->dense_small.ptr;
  uint32_t* binbase = btot + kSmallStride;
  uint32_t* tbtot = binbase + kSmallStride;
  uint32_t* tbinbase = tbtot + kSmallStride;
  uint32_t* fcur = tbinbase + kSmallStride;
  uint64_t* ocnt = (uint64_t*)((uint32_t*)ctx->dense_small.ptr + kSmallWords);
  uint64_t* obase = ocnt + kMaxSlices + 64;
  uint32_t* ent_d = (uint32_t*)ctx->dense_ent.ptr;
  T* ent_v = (T*)((char*)ctx->dense_ent.ptr + (((size_t)k * 4 + 255) & ~(size_t)255));
  const size_t fcap = (size_t)d.NBf << d.bsf;
  uint32_t* fent_d = (uint32_t*)ctx->dense_fine.ptr;
  T* fent_v = (T*)((char*)ctx->dense_fine.ptr + ((fcap * 4 + 255) & ~(size_t)255));
  T* src = (T*)ctx->dense_src.ptr;
  uint32_t* dpos = (uint32_t*)((char*)ctx->dense_src.ptr + (((size_t)src_cap * sizeof(T) + 255) & ~(size_t)255));
  if (d.NBt) {
    // tail marks grouped by tail region (the tail bitmap is zero here: every
    // launch form leaves it clear)
    rz_status st = bucket_marks(ctx, R, k, sZ, k, bits, nullptr,
                                stats ? (uint64_t*)(ctrl + kCtrlStatTail) : (uint64_t*)nullptr, s);
    if (st != RZ_OK) return st;
  }
  uint32_t* off = cnt + (size_t)d.nt * d.NB;
  hipLaunchKernelGGL((rz_dense_count<WG, IPT>), dim3(d.nt), dim3(WG), 0, s, R, k, sZ, d.L, bits, d.NB, d.bs, cnt,
                     ocnt, fcur, d.NBf);
  hipLaunchKernelGGL(rz_dense_colscan, dim3(d.NB / 32), dim3(1024), 0, s, cnt, d.nt, d.NB, off, btot);
  hipLaunchKernelGGL(rz_dense_topscan, dim3(1), dim3(1024), 0, s, btot, d.NB, binbase, ocnt, d.nt, obase, ctrl,
                     stats);
  hipLaunchKernelGGL((rz_dense_entries<T, WG, IPT>), dim3(d.nt), dim3(WG), lds_ent, s, A, R, k, sZ, d.L, bits, d.NB,
                     d.bs, off, binbase, obase, ent_d, ent_v, src, dpos, src_cap);
  hipLaunchKernelGGL((rz_dense_refine<T, WG, IPT>), dim3((k + S - 1) / S), dim3(WG), lds_ent, s, ent_d, ent_v,
                     binbase + d.NB, fent_d, fent_v, fcur, d.NBf, d.bsf, ctrl);
  // orphan pairs (dpos[t], src[t]), read in rank order, into the same fine bins
  hipLaunchKernelGGL((rz_dense_refine<T, WG, IPT>), dim3((src_cap + S - 1) / S), dim3(WG), lds_ent, s, dpos, src,
                     (const uint32_t*)(obase + d.nt + 1), fent_d, fent_v, fcur, d.NBf, d.bsf, ctrl);
  hipLaunchKernelGGL(rz_dense_scatter<T>, dim3(d.NBf * d.G), dim3(1024), 0, s, A, fent_d, fent_v, fcur, d.NBf, d.bsf,
                     d.G, sZ);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

bool dense_bucket_marks(const rz_tile_config& cfg, int64_t k, uint32_t flags) {
  return cfg.variant == 5 && cfg.chunk == 2 && k <= ((int64_t)kCoarseMax << kTailBinLog) && !(flags & RZ_SORTED);
}

// The dense form's bins and tiles for (cfg, n, k); RZ_EINVAL when the shape
// cannot be launched (bin counts above their limits, LDS above the device's,
// or a workgroup shape with no kernel).  No side effect, so a caller can
// check before it launches anything.
static rz_status dense_shape(const rz_ctx* ctx, const rz_tile_config& cfg, uint32_t esz, int64_t n, int64_t k,
                             bool bucket_marks, DenseShape* out) {
  DenseShape& d = *out;
  const int64_t sZ = n - k;
  slice_shape(k, cfg, kMaxSlices, &d.L, &d.nt);
  d.bs = bin_log(sZ, kCoarseTarget);
  d.NB = (uint32_t)((sZ + ((int64_t)1 << d.bs) - 1) >> d.bs);
  d.NB = d.NB < 32 ? 32 : (d.NB + 31u) & ~31u;
  int lg = cfg.slice_steps > 0 ? cfg.slice_steps : 10;
  if (lg > 12) lg = 12;
  if (lg < 3) lg = 3;
  d.bsf = bin_log(sZ, (int64_t)1 << lg);
  d.NBf = (uint32_t)((sZ + ((int64_t)1 << d.bsf) - 1) >> d.bsf);
  d.NBf = (d.NBf + 7u) & ~7u;
  d.NBt = 0;
  if (bucket_marks) {
    d.NBt = (uint32_t)((k + ((int64_t)1 << kTailBinLog) - 1) >> kTailBinLog);
    d.NBt = (d.NBt + 31u) & ~31u;
  }
  // scatter: every XCD holds about two 1024-thread workgroups per CU, so
  // G = CUs per XCD keeps about two fine bins of each XCD in flight
  d.G = (uint32_t)(ctx->cu_count / 8 > 0 ? ctx->cu_count / 8 : 1);
  if (d.NB > (uint32_t)kCoarseMax || d.NBf > (uint32_t)kFineMax || d.NBt > (uint32_t)kCoarseMax) return RZ_EINVAL;
  const size_t lds = (size_t)cfg.wg_size * cfg.items * (esz + 6) + 48 * 1024;  // + static LDS (refine)
  if (lds > (size_t)ctx->lds_max) return RZ_EINVAL;
  const bool kernel = (cfg.wg_size == 256 || cfg.wg_size == 512 || cfg.wg_size == 1024) &&
                      (cfg.items == 4 || cfg.items == 8);
  return kernel ? RZ_OK : RZ_EINVAL;
}

bool dense_shape_ok(const rz_ctx* ctx, const rz_tile_config& cfg, uint32_t esz, int64_t n, int64_t k, uint32_t flags) {
  DenseShape d;
  return dense_shape(ctx, cfg, esz, n, k, dense_bucket_marks(cfg, k, flags), &d) == RZ_OK;
}

template <typename T>
rz_status dense_moves(rz_ctx* ctx, T* A, int64_t n, const int32_t* R, int64_t k, const rz_tile_config& cfg,
                      uint32_t* bits, uint32_t* ctrl, int stats, bool bucket_marks, hipStream_t s) {
  const int64_t sZ = n - k;
  DenseShape d;
  rz_status st = dense_shape(ctx, cfg, (uint32_t)sizeof(T), n, k, bucket_marks, &d);
  if (st != RZ_OK) return st;
  const uint32_t src_cap = (uint32_t)(k / 2 + 1);
  const uint32_t nbmax = d.NB > d.NBt ? d.NB : d.NBt;
  const size_t fcap = (size_t)d.NBf << d.bsf;
  if ((st = ensure(ctx, ctx->dense_cnt, (size_t)d.nt * nbmax * 4 * 2, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->dense_small, kSmallBytes, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->dense_ent, (((size_t)k * 4 + 255) & ~(size_t)255) + (size_t)k * sizeof(T), s)) != RZ_OK)
    return st;
  if ((st = ensure(ctx, ctx->dense_fine, ((fcap * 4 + 255) & ~(size_t)255) + fcap * sizeof(T), s)) != RZ_OK)
    return st;
  if ((st = ensure(ctx, ctx->dense_src, (((size_t)src_cap * sizeof(T) + 255) & ~(size_t)255) + (size_t)src_cap * 4,
                   s)) != RZ_OK)
    return st;
#define RZ_DENSE_CASE(W, I) \
  if (cfg.wg_size == W && cfg.items == I) return dense_launch<T, W, I>(ctx, A, R, (uint32_t)k, (uint32_t)sZ, d, bits, ctrl, stats, src_cap, s);
  RZ_DENSE_CASE(256, 4) RZ_DENSE_CASE(256, 8) RZ_DENSE_CASE(512, 4) RZ_DENSE_CASE(512, 8)
  RZ_DENSE_CASE(1024, 4) RZ_DENSE_CASE(1024, 8)
#undef RZ_DENSE_CASE
  return RZ_EINVAL;
}

template rz_status dense_moves<uint32_t>(rz_ctx*, uint32_t*, int64_t, const int32_t*, int64_t,
                                         const rz_tile_config&, uint32_t*, uint32_t*, int, bool, hipStream_t);
template rz_status dense_moves<uint64_t>(rz_ctx*, uint64_t*, int64_t, const int32_t*, int64_t,
                                         const rz_tile_config&, uint32_t*, uint32_t*, int, bool, hipStream_t);
template rz_status dense_moves<uint4>(rz_ctx*, uint4*, int64_t, const int32_t*, int64_t, const rz_tile_config&,
                                      uint32_t*, uint32_t*, int, bool, hipStream_t);

}  // namespace rz
