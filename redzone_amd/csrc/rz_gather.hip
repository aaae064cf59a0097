// rz_gather.hip -- add+remove, and the in-place gather by a preserve list.
//
// Preserve-list gather (paper p.14, eq. 3): A'[i] = A[P[i]] for i < p, P
// strictly ascending.  Since P[i] >= i, items only move down.  Tile t (outputs
// [t*TILE, (t+1)*TILE)) reads its items into registers, waits until the reads
// have returned, publishes "read" in its status word, and writes only after
// every earlier tile has published: tile t's outputs can overlap only the
// inputs of tiles t' <= t (an input position P[i'] inside tile t's outputs
// means i' <= P[i'] < (t+1)*TILE), so no item is overwritten before it is
// read.  The wait is a decoupled look-back over the status words (flag 1 =
// this tile has read, flag 2 = this tile and every earlier tile have read),
// so a tile walks back only to the nearest tile with flag 2.  Status words
// carry a 30-bit call epoch; the last workgroup to leave advances it.
#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

constexpr uint64_t kGatherRead = 1ull, kGatherAllRead = 2ull;

__device__ __forceinline__ uint64_t gather_status(uint64_t flag, uint32_t epoch) {
  return (flag << 62) | ((uint64_t)(epoch & 0x3fffffffu) << 32);
}

template <typename T, int WG, int IPT>
__global__ __launch_bounds__(WG) void rz_gather_kernel(T* __restrict__ A, uint32_t n, const int32_t* __restrict__ P,
                                                       uint32_t p, uint64_t* __restrict__ status,
                                                       uint32_t* __restrict__ ctrl) {
  const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint32_t t = blockIdx.x;
  const uint32_t base = t * (uint32_t)(WG * IPT);
  const uint32_t epoch = ld_relaxed(&ctrl[kCtrlGatherEpoch]);
  const uint32_t ep = epoch & 0x3fffffffu;
  T x[IPT];
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * WG + tid;
    if (i < p) {
      const uint32_t src = (uint32_t)P[i];
      x[j] = A[src < n ? src : i];  // an index out of range (invalid list) reads in bounds
    }
  }
  wait_all_memory();  // every read of this tile has returned
  __syncthreads();
  if (wave == 0) {
    if (t == 0) {
      if (lane == 0) st_relaxed(&status[0], gather_status(kGatherAllRead, ep));
    } else {
      if (lane == 0) st_relaxed(&status[t], gather_status(kGatherRead, ep));
      int64_t jb = (int64_t)t - 1;
      uint32_t spins = 0;
      while (true) {
        const int64_t q = jb - lane;
        const uint64_t st = q >= 0 ? ld_relaxed(&status[q]) : gather_status(kGatherAllRead, ep);
        uint32_t flag = (((uint32_t)(st >> 32) & 0x3fffffffu) == ep) ? (uint32_t)(st >> 62) : 0u;
        if (q < 0) flag = (uint32_t)kGatherAllRead;
        const uint64_t all_m = ballot(flag == kGatherAllRead);
        const uint64_t none_m = ballot(flag == 0u);
        // lanes up to and including the nearest "all read" must all have read
        const uint64_t upto = all_m ? ((all_m & (~all_m + 1)) << 1) - 1 : ~0ull;
        if (none_m & upto) {
          if (++spins > (1u << 24)) {  // give up (the grid is not making progress)
            if (lane == 0) or_relaxed(&ctrl[kCtrlError], kErrLookbackTimeout);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        if (all_m) break;
        jb -= kWave;
      }
      if (lane == 0) st_relaxed(&status[t], gather_status(kGatherAllRead, ep));
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < IPT; ++j) {
    const uint32_t i = base + j * WG + tid;
    if (i < p) A[i] = x[j];
  }
  // the last workgroup to leave (all of them have read the epoch) advances it
  if (tid == 0 && add_relaxed(&ctrl[kCtrlGatherExit], 1u) == gridDim.x - 1u) {
    st_relaxed(&ctrl[kCtrlGatherExit], 0u);
    st_relaxed(&ctrl[kCtrlGatherEpoch], epoch + 1u);
  }
}

template <typename T, int IPT>
static rz_status gather_launch(rz_ctx* ctx, T* A, uint32_t n, const int32_t* P, uint32_t p, hipStream_t s) {
  constexpr int WG = 256;
  const uint32_t tiles = (p + WG * IPT - 1) / (WG * IPT);
  rz_status st;
  if ((st = ensure(ctx, ctx->gstatus, (size_t)tiles * 8, s)) != RZ_OK) return st;
  hipLaunchKernelGGL((rz_gather_kernel<T, WG, IPT>), dim3(tiles), dim3(WG), 0, s, A, n, P, p,
                     (uint64_t*)ctx->gstatus.ptr, (uint32_t*)ctx->ctrl.ptr);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
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
      case 4: st = gather_launch<uint32_t, 8>(ctx, (uint32_t*)d_data, (uint32_t)n, d_keep, (uint32_t)p, s); break;
      case 8: st = gather_launch<uint64_t, 8>(ctx, (uint64_t*)d_data, (uint32_t)n, d_keep, (uint32_t)p, s); break;
      default: st = gather_launch<uint4, 4>(ctx, (uint4*)d_data, (uint32_t)n, d_keep, (uint32_t)p, s); break;
    }
  }
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
