// rz_gather.hip -- add+remove, and the in-place gather by a preserve list.
//
// Preserve-list gather (paper p.14, eq. 3): A'[i] = A[P[i]] for i < p, P
// strictly ascending.  Two forms:
//  - one pass, in place (p >= kGatherOnePassMin): tiles of outputs in
//    blockIdx order; a tile writes once every earlier tile has read its
//    sources (loaded flags, summarised per group of 64 tiles), the same guard
//    as the partition's in-place compaction.  Streams P, A[P] and the output
//    once: 4.25 TB/s at p = 0.99 n on MI355X.
//  - two passes (small p): gather into context scratch with no ordering
//    between workgroups, then a device copy back to the front of the list.
//    Below a few million items the one-pass form's flag traffic and serial
//    wait cost more than the copy.
#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

constexpr uint32_t kGatherOnePassMin = 1u << 22;

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_gather_kernel(const T* __restrict__ A, uint32_t n,
                                                       const int32_t* __restrict__ P, uint32_t p,
                                                       T* __restrict__ out) {
  const uint32_t base = blockIdx.x * (uint32_t)(WG * IPT) + threadIdx.x;
  uint32_t src[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * WG;
    src[j] = i < p ? (uint32_t)P[i] : 0u;
  }
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * WG;
    if (i < p) out[i] = A[src[j] < n ? src[j] : i];  // an index out of range (invalid list) reads in bounds
  }
}

// One pass, in place: tile t gathers outputs [t*TILE, (t+1)*TILE) into
// registers, sets its loaded flag (and its group word once all 64 tiles of
// its group have loaded), waits until every earlier tile has loaded, and
// writes its outputs.  The sources of a tile lie at or after its first output
// (P ascends, P[i] >= i), so only earlier tiles' sources can sit in its output
// range, and later tiles never write below it.  Flags carry the call's tag,
// the epoch word plus one; the last tile to leave (every tile has read the
// epoch by then) stores the tag as the new epoch, so no flag is ever cleared
// and no separate launch advances the epoch.
template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_gather_inplace(T* __restrict__ A, uint32_t n,
                                                        const int32_t* __restrict__ P, uint32_t p,
                                                        uint32_t* __restrict__ loaded, uint32_t* __restrict__ gcnt,
                                                        uint32_t* __restrict__ gdone, uint32_t ntiles,
                                                        uint32_t* __restrict__ ctrl) {
  constexpr uint32_t TILE = WG * IPT;
  const uint32_t t = blockIdx.x, i0 = t * TILE;
  const uint32_t e0 = ld_relaxed(&ctrl[kCtrlEpoch]);
  const uint32_t epoch = (e0 + 1u) & 0x3fffffffu;  // this call's tag
  uint32_t src[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = i0 + j * WG + threadIdx.x;
    src[j] = i < p ? (uint32_t)P[i] : 0u;
  }
  T v[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = i0 + j * WG + threadIdx.x;
    v[j] = T();
    if (i < p) v[j] = A[src[j] < n ? src[j] : i];  // an index out of range (invalid list) reads in bounds
  }
  wait_all_memory();
  __syncthreads();
  if (threadIdx.x == 0) {
    st_relaxed(&loaded[t], epoch);
    const uint32_t g = t >> 6, members = min(64u, ntiles - (g << 6));
    if (add_relaxed(&gcnt[g], 1u) == members - 1u) {
      st_relaxed(&gcnt[g], 0u);
      st_relaxed(&gdone[g], epoch);
    }
  }
  if (threadIdx.x < kWave && t != 0) {
    // every earlier tile has loaded: whole groups by their word, this tile's
    // own group tile by tile; up to 8 words per lane per round trip
    const int lane = lane_id();
    const uint32_t g1 = t >> 6, tb = g1 << 6;
    const uint32_t total = g1 + (t - tb);
    uint32_t spins = 0, pause = 1;
    for (uint32_t c0 = 0; c0 < total;) {
      bool ok = true;
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const uint32_t c = c0 + (uint32_t)(r * kWave + lane);
        if (c < total) ok &= (c < g1 ? ld_relaxed(&gdone[c]) : ld_relaxed(&loaded[tb + (c - g1)])) == epoch;
      }
      if (ballot(!ok)) {
        if (++spins > spin_cap(ctrl, 1u << 22)) {
          if (lane == 0) or_relaxed(&ctrl[kCtrlError], kErrLookbackTimeout);
          break;
        }
        for (uint32_t i = 0; i < pause; ++i) __builtin_amdgcn_s_sleep(1);
        if (pause < 16u) pause <<= 1;
        continue;
      }
      c0 += 8 * kWave;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = i0 + j * WG + threadIdx.x;
    if (i < p) A[i] = v[j];
  }
  if (threadIdx.x == 0 && add_relaxed(&ctrl[kCtrlGatherExit], 1u) == ntiles - 1u) {
    st_relaxed(&ctrl[kCtrlGatherExit], 0u);
    st_relaxed(&ctrl[kCtrlEpoch], e0 + 1u);
  }
}

template <typename T, int IPT>
static rz_status gather_inplace_launch(rz_ctx* ctx, T* A, uint32_t n, const int32_t* P, uint32_t p, hipStream_t s) {
  constexpr int WG = 1024;
  const uint32_t tiles = (p + WG * IPT - 1) / (WG * IPT);
  const uint32_t groups = (tiles + 63) / 64;
  rz_status st;
  if ((st = ensure(ctx, ctx->ploaded, (size_t)tiles * 4, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->pgcnt, (size_t)groups * 4, s)) != RZ_OK) return st;
  if ((st = ensure(ctx, ctx->pgdone, (size_t)groups * 4, s)) != RZ_OK) return st;
  uint32_t* ctrl = (uint32_t*)ctx->ctrl.ptr;
  hipLaunchKernelGGL((rz_gather_inplace<T, WG, IPT>), dim3(tiles), dim3(WG), 0, s, A, n, P, p,
                     (uint32_t*)ctx->ploaded.ptr, (uint32_t*)ctx->pgcnt.ptr, (uint32_t*)ctx->pgdone.ptr, tiles, ctrl);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

template <typename T, int IPT>
static rz_status gather_launch(rz_ctx* ctx, T* A, uint32_t n, const int32_t* P, uint32_t p, hipStream_t s) {
  constexpr int WG = 256;
  rz_status st;
  if ((st = ensure(ctx, ctx->gstatus, (size_t)p * sizeof(T), s)) != RZ_OK) return st;
  T* out = (T*)ctx->gstatus.ptr;
  const uint32_t tiles = (p + WG * IPT - 1) / (WG * IPT);
  hipLaunchKernelGGL((rz_gather_kernel<T, WG, IPT>), dim3(tiles), dim3(WG), 0, s, A, n, P, p, out);
  RZ_HIP(hipGetLastError());
  RZ_HIP(hipMemcpyAsync(A, out, (size_t)p * sizeof(T), hipMemcpyDeviceToDevice, s));
  return RZ_OK;
}

template <typename T, int IPT>
static rz_status gather_form(rz_ctx* ctx, void* A, int64_t n, const int32_t* P, int64_t p, hipStream_t s) {
  if ((uint32_t)p >= kGatherOnePassMin) return gather_inplace_launch<T, IPT>(ctx, (T*)A, (uint32_t)n, P, (uint32_t)p, s);
  return gather_launch<T, IPT>(ctx, (T*)A, (uint32_t)n, P, (uint32_t)p, s);
}

}  // namespace rz

using namespace rz;

extern "C" {

rz_status rz_gather_preserve(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, const int32_t* d_keep,
                             int64_t p, uint32_t flags, void* stream) {
  if (!ctx || n < 0 || p < 0 || p > n || n >= (1ll << 31) || esz_index(elem_bytes) < 0) return RZ_EINVAL;
  if (p == 0) return RZ_OK;
  if (!d_data || !d_keep) return RZ_EINVAL;
  if (((uintptr_t)d_data % (elem_bytes > 8 ? 16 : elem_bytes)) != 0 || ((uintptr_t)d_keep % 4) != 0) return RZ_EINVAL;
  std::lock_guard<std::mutex> lock(ctx->mu);
  int prev = -1;
  (void)hipGetDevice(&prev);
  if (prev != ctx->device) (void)hipSetDevice(ctx->device);
  hipStream_t s = (hipStream_t)stream;
  rz_status st = RZ_OK;
  if (flags & RZ_VALIDATE) st = launch_validate(ctx, d_keep, n, p, RZ_SORTED, s);
  if (st == RZ_OK) {
    switch (elem_bytes) {
      case 4: st = gather_form<uint32_t, 8>(ctx, d_data, n, d_keep, p, s); break;
      case 8: st = gather_form<uint64_t, 8>(ctx, d_data, n, d_keep, p, s); break;
      default: st = gather_form<uint4, 4>(ctx, d_data, n, d_keep, p, s); break;
    }
  }
  if (st == RZ_OK) st = finish_call(ctx, flags, s);
  if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
  return st;
}

rz_status rz_add_remove(rz_ctx* ctx, void* d_data, int64_t n, int64_t capacity, uint32_t elem_bytes,
                        const void* d_add, int64_t g, int32_t* d_removals, int64_t k, uint32_t flags, void* stream) {
  if (!ctx || n < 0 || g < 0 || capacity < n + g || esz_index(elem_bytes) < 0) return RZ_EINVAL;
  if (g > 0 && (!d_add || !d_data)) return RZ_EINVAL;
  const int64_t m = n + g;
  if (m >= (1ll << 31) || k < 0 || k > m) return RZ_EINVAL;
  if (g > 0) {
    const char* add = (const char*)d_add;
    char* dst = (char*)d_data + (size_t)n * elem_bytes;
    const size_t bytes = (size_t)g * elem_bytes;
    if (add < dst + bytes && dst < add + bytes) return RZ_EINVAL;  // the added items may not overlap their slot
    int prev = -1;
    (void)hipGetDevice(&prev);
    if (prev != ctx->device) (void)hipSetDevice(ctx->device);
    const hipError_t e = hipMemcpyAsync(dst, add, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
    if (prev >= 0 && prev != ctx->device) (void)hipSetDevice(prev);
    if (e != hipSuccess) return RZ_EHIP;
  }
  return rz_remove(ctx, d_data, m, elem_bytes, d_removals, k, flags, stream);
}

}  // extern "C"
