// rz_gather.hip -- add+remove, and the in-place gather by a preserve list.
//
// Preserve-list gather (paper p.14, eq. 3): A'[i] = A[P[i]] for i < p, P
// strictly ascending.  Done in two streaming passes with no ordering between
// workgroups: the gather reads P and A[P[i]] (nearly sequential addresses, as
// P ascends) into context scratch, then a device copy writes the p items back
// to the front of the list.  (A single in-place pass has to make every tile
// wait until all earlier tiles have read their items; measured on MI355X that
// ordering held the pass to about 440 GB/s, a third of what the two passes'
// extra traffic costs.)
#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

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
