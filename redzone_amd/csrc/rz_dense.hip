// rz_dense.hip -- region-bucketed Redzone for dense removal lists (variant 5).
//
// A scattered 4-byte write costs one read-modify-write of a DRAM burst, so the
// move step of the plain forms runs at about 25 G holes/s whatever the
// density.  When holes are dense (k/n of several percent and up) most bursts
// hold several holes, and writing them while the burst's line sits in one
// L2 turns many read-modify-writes into one.  This form therefore regroups
// the writes by destination before issuing them:
//
//   mark        tail bitmap (rz_mark_tail, as in the other forms)
//   count       tile j of L pairs: per destination bin, the number of head
//               writes (moves and dest orphans), and the tile's dest / source
//               orphan counts.  Bins are 2^BS consecutive positions of [0, n-k).
//   colscan     per bin, exclusive prefix of the counts over tiles
//   topscan     bin bases; per tile, exclusive dest / source orphan ranks
//   entries     tile j again: each head write goes to its bin's run in
//               (d, value) arrays; a dest orphan carries its rank t instead of
//               a value (top bit of d set); the t-th source orphan's value goes
//               to src[t].  Orphan ranks follow i, so pairing stays canonical
//               (redzone.hxx:834-934).
//   scatter     each bin is split among G workgroups laid out so that all of
//               them share one XCD (workgroups are dispatched to the 8 XCDs
//               round robin); A[d] = value, or src[t] for a dest orphan.
//
// The counts and offsets are recomputed on every call, so no scratch needs
// clearing except the tail bitmap, which the entries kernel clears as it
// reads it.  Distinct removals give at most one head write per position, so
// every bin run and every index stays inside its buffer; with an invalid
// list the two passes still classify identically, so indices stay in bounds
// and the orphan-count mismatch is flagged.
#include "rz_redzone_dev.hpp"

namespace rz {

constexpr int kDenseMaxBins = 1024;
constexpr uint32_t kDenseOrphan = 0x80000000u;

__device__ __forceinline__ uint32_t classify(uint32_t d, bool dead, uint32_t sZ) {
  const bool dst_ok = d < sZ;
  return dead ? (dst_ok ? kDstOrphan : kNone) : (dst_ok ? kMove : kSrcOrphan);
}

// ---- count ----------------------------------------------------------------
template <int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_dense_count(const int32_t* __restrict__ R, uint32_t k, uint32_t sZ,
                                                      uint32_t L, const uint32_t* __restrict__ bits, uint32_t NB,
                                                      uint32_t bs, uint32_t* __restrict__ cnt,
                                                      uint64_t* __restrict__ ocnt) {
  __shared__ uint32_t hist[kDenseMaxBins];
  __shared__ uint32_t s_nd, s_ns;
  const uint32_t tid = threadIdx.x, j = blockIdx.x;
  const uint32_t a = j * L, b = min(a + L, k);
  for (uint32_t q = tid; q < NB; q += WG) hist[q] = 0u;
  if (tid == 0) { s_nd = 0u; s_ns = 0u; }
  __syncthreads();
  uint32_t nd = 0, ns = 0;
  for (uint32_t base = a; base < b; base += WG * IPT) {
    uint32_t d[IPT], wv[IPT];
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t p = base + q * WG + tid;
      if (p < b) { d[q] = (uint32_t)R[p]; wv[q] = bits[p >> 5]; }
    }
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t p = base + q * WG + tid;
      if (p >= b) continue;
      const uint32_t cls = classify(d[q], (wv[q] >> (p & 31)) & 1u, sZ);
      if (cls == kMove || cls == kDstOrphan) atomicAdd(&hist[d[q] >> bs], 1u);
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
    if (td != ts) or_relaxed(&ctrl[kCtrlError], kErrOrphanMismatch);
    if (stats) st_relaxed((uint64_t*)&ctrl[kCtrlStatOrphans], (uint64_t)min(td, ts));
  }
}

// ---- entries ----------------------------------------------------------------
template <typename T> __device__ __forceinline__ T rank_value(uint32_t t);
template <> __device__ __forceinline__ uint32_t rank_value<uint32_t>(uint32_t t) { return t; }
template <> __device__ __forceinline__ uint64_t rank_value<uint64_t>(uint32_t t) { return t; }
template <> __device__ __forceinline__ uint4 rank_value<uint4>(uint32_t t) { return make_uint4(t, 0u, 0u, 0u); }
template <typename T> __device__ __forceinline__ uint32_t value_rank(const T& v);
template <> __device__ __forceinline__ uint32_t value_rank<uint32_t>(const uint32_t& v) { return v; }
template <> __device__ __forceinline__ uint32_t value_rank<uint64_t>(const uint64_t& v) { return (uint32_t)v; }
template <> __device__ __forceinline__ uint32_t value_rank<uint4>(const uint4& v) { return v.x; }

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_dense_entries(const T* __restrict__ A, const int32_t* __restrict__ R,
                                                        uint32_t k, uint32_t sZ, uint32_t L,
                                                        uint32_t* __restrict__ bits, uint32_t NB, uint32_t bs,
                                                        const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ binbase,
                                                        const uint64_t* __restrict__ obase,
                                                        uint32_t* __restrict__ ent_d, T* __restrict__ ent_v,
                                                        T* __restrict__ src, uint32_t src_cap) {
  constexpr int NW = WG / kWave;
  constexpr int E = MoveShared<WG, IPT>::E;
  constexpr int PER = (E + kWave - 1) / kWave;
  __shared__ MoveShared<WG, IPT> sh;
  __shared__ uint32_t cur[kDenseMaxBins];
  uint32_t* s_off = sh.off;
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint32_t j = blockIdx.x;
  const uint32_t a = j * L, b = min(a + L, k);
  const T* __restrict__ Z = A + sZ;
  for (uint32_t q = tid; q < NB; q += WG) cur[q] = binbase[q] + off[(size_t)j * NB + q];
  const uint64_t ob = obase[j];
  const uint32_t xd = (uint32_t)ob, xs = (uint32_t)(ob >> 32);
  __syncthreads();
  uint32_t carry_s = 0, carry_d = 0;
  for (uint32_t base = a; base < b; base += WG * IPT) {
    uint32_t d[IPT], wv[IPT], info[IPT];
    T z[IPT];
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t p = base + q * WG + tid;
      if (p < b) {
        d[q] = (uint32_t)R[p];
        z[q] = Z[p];
        wv[q] = bits[p >> 5];
      } else {
        d[q] = sZ;
        wv[q] = 0xffffffffu;
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
    __syncthreads();
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
      const uint32_t cls = info[q] & 0xffu;
      const uint32_t o = s_off[q * NW + wave];
      if (cls == kSrcOrphan) {
        const uint32_t t = xs + carry_s + (o & 0xffffu) + ((info[q] >> 8) & 0xffu);
        if (t < src_cap) src[t] = z[q];
      } else if (cls != kNone) {
        const uint32_t slot = atomicAdd(&cur[d[q] >> bs], 1u);
        const bool mv = cls == kMove;
        const uint32_t t = xd + carry_d + (o >> 16) + (info[q] >> 16);
        ent_d[slot] = mv ? d[q] : (d[q] | kDenseOrphan);
        ent_v[slot] = mv ? z[q] : rank_value<T>(t);
      }
    }
    const uint32_t tot = s_off[E];
    carry_s += tot & 0xffffu;
    carry_d += tot >> 16;
    __syncthreads();
  }
  // leave the tail bitmap zero for the next call (tiles are word aligned)
  for (uint32_t w = (a >> 5) + tid; w < ((b + 31u) >> 5); w += WG) bits[w] = 0u;
}

// ---- scatter ----------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(1024) void rz_dense_scatter(T* __restrict__ A, const uint32_t* __restrict__ ent_d,
                                                        const T* __restrict__ ent_v,
                                                        const uint32_t* __restrict__ binbase, uint32_t G,
                                                        const T* __restrict__ src, uint32_t src_cap, uint32_t sZ) {
  const uint32_t w = blockIdx.x;
  const uint32_t x = w & 7u, rest = w >> 3;
  const uint32_t g = rest % G, bin = (rest / G) * 8u + x;
  const uint32_t s = binbase[bin], e = binbase[bin + 1];
  const uint32_t per = (e - s + G - 1) / G;
  const uint32_t lo = s + g * per, hi = min(lo + per, e);
  for (uint32_t q = lo + threadIdx.x; q < hi; q += 1024u) {
    uint32_t d = ent_d[q];
    T v = ent_v[q];
    if (d & kDenseOrphan) {
      const uint32_t t = value_rank<T>(v);
      if (t >= src_cap) continue;
      v = src[t];
      d &= ~kDenseOrphan;
    }
    if (d < sZ) A[d] = v;
  }
}

// ---------------------------------------------------------------------------
// launch
// ---------------------------------------------------------------------------
struct DenseShape {
  uint32_t L, nt, NB, bs, G;
};

// Bins: NB (a multiple of 32, at most 1024) bins of 2^bs positions covering
// [0, sZ); 2^slice_steps is the target bin count.  Tiles: as the move slices.
static bool dense_shape(int64_t n, int64_t k, const rz_tile_config& cfg, DenseShape* sh) {
  const int64_t sZ = n - k;
  int lg = cfg.slice_steps > 0 ? cfg.slice_steps : 9;
  if (lg > 10) lg = 10;
  if (lg < 5) lg = 5;
  uint32_t bs = 10;
  while (((sZ + ((int64_t)1 << bs) - 1) >> bs) > ((int64_t)1 << lg)) ++bs;
  uint32_t nb = (uint32_t)((sZ + ((int64_t)1 << bs) - 1) >> bs);
  nb = (nb + 31u) & ~31u;
  if (nb < 32) nb = 32;
  if (nb > (uint32_t)kDenseMaxBins) return false;
  uint32_t L, nt;
  slice_shape(k, cfg, kMaxSlices, &L, &nt);
  uint32_t G = 2048u / nb;
  if (G < 1) G = 1;
  *sh = DenseShape{L, nt, nb, bs, G};
  return true;
}

template <typename T, int WG, int IPT>
static void dense_launch(rz_ctx* ctx, T* A, const int32_t* R, uint32_t k, uint32_t sZ, const DenseShape& d,
                         uint32_t* bits, uint32_t* ctrl, int stats, uint32_t src_cap, hipStream_t s) {
  uint32_t* cnt = (uint32_t*)ctx->dense_cnt.ptr;
  uint32_t* off = cnt + (size_t)d.nt * d.NB;
  uint32_t* btot = (uint32_t*)ctx->dense_small.ptr;
  uint32_t* binbase = btot + kDenseMaxBins;
  uint64_t* ocnt = (uint64_t*)(binbase + kDenseMaxBins + 64);
  uint64_t* obase = ocnt + kMaxSlices;
  uint32_t* ent_d = (uint32_t*)ctx->dense_ent.ptr;
  T* ent_v = (T*)((char*)ctx->dense_ent.ptr + (((size_t)k * 4 + 255) & ~(size_t)255));
  T* src = (T*)ctx->dense_src.ptr;
  hipLaunchKernelGGL((rz_dense_count<WG, IPT>), dim3(d.nt), dim3(WG), 0, s, R, k, sZ, d.L, bits, d.NB, d.bs, cnt,
                     ocnt);
  hipLaunchKernelGGL(rz_dense_colscan, dim3(d.NB / 32), dim3(1024), 0, s, cnt, d.nt, d.NB, off, btot);
  hipLaunchKernelGGL(rz_dense_topscan, dim3(1), dim3(1024), 0, s, btot, d.NB, binbase, ocnt, d.nt, obase, ctrl,
                     stats);
  hipLaunchKernelGGL((rz_dense_entries<T, WG, IPT>), dim3(d.nt), dim3(WG), 0, s, A, R, k, sZ, d.L, bits, d.NB, d.bs,
                     off, binbase, obase, ent_d, ent_v, src, src_cap);
  hipLaunchKernelGGL(rz_dense_scatter<T>, dim3(d.NB * d.G), dim3(1024), 0, s, A, ent_d, ent_v, binbase, d.G, src,
                     src_cap, sZ);
}

template <typename T>
rz_status dense_moves(rz_ctx* ctx, T* A, int64_t n, const int32_t* R, int64_t k, const rz_tile_config& cfg,
                      uint32_t* bits, uint32_t* ctrl, int stats, hipStream_t s) {
  DenseShape d;
  if (!dense_shape(n, k, cfg, &d)) return RZ_EINVAL;
  const uint32_t src_cap = (uint32_t)(k / 2 + 1);
  rz_status st;
  if ((st = ensure(ctx, ctx->dense_cnt, (size_t)d.nt * d.NB * 4 * 2, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->dense_small, (size_t)(2 * kDenseMaxBins + 64) * 4 + (size_t)2 * (kMaxSlices + 1) * 8,
                   s)) != RZ_OK)
    return st;
  if ((st = ensure(ctx, ctx->dense_ent, (((size_t)k * 4 + 255) & ~(size_t)255) + (size_t)k * sizeof(T), s)) != RZ_OK)
    return st;
  if ((st = ensure(ctx, ctx->dense_src, (size_t)src_cap * sizeof(T), s)) != RZ_OK) return st;
  const uint32_t sZ = (uint32_t)(n - k);
#define RZ_DENSE_CASE(W, I)                                                                             \
  if (cfg.wg_size == W && cfg.items == I) {                                                             \
    dense_launch<T, W, I>(ctx, A, R, (uint32_t)k, sZ, d, bits, ctrl, stats, src_cap, s);                \
    RZ_HIP(hipGetLastError());                                                                          \
    return RZ_OK;                                                                                       \
  }
  RZ_DENSE_CASE(256, 4) RZ_DENSE_CASE(256, 8) RZ_DENSE_CASE(512, 4) RZ_DENSE_CASE(512, 8)
  RZ_DENSE_CASE(1024, 4) RZ_DENSE_CASE(1024, 8)
#undef RZ_DENSE_CASE
  return RZ_EINVAL;
}

template rz_status dense_moves<uint32_t>(rz_ctx*, uint32_t*, int64_t, const int32_t*, int64_t,
                                         const rz_tile_config&, uint32_t*, uint32_t*, int, hipStream_t);
template rz_status dense_moves<uint64_t>(rz_ctx*, uint64_t*, int64_t, const int32_t*, int64_t,
                                         const rz_tile_config&, uint32_t*, uint32_t*, int, hipStream_t);
template rz_status dense_moves<uint4>(rz_ctx*, uint4*, int64_t, const int32_t*, int64_t, const rz_tile_config&,
                                      uint32_t*, uint32_t*, int, hipStream_t);

}  // namespace rz
