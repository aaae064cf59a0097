// rz_gen.hip -- device generators of synthetic inputs (autotuner, mixed
// streams too large to build on the host).  Integer-exact: they produce the
// same values as redzone_amd/workloads.py (uniform_removals, unique_u32_payload,
// u64_payload; 16-byte records carry a hash in their fourth word where the
// Python generator draws a float).
#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

struct PermKeys {
  uint32_t key[4];
  uint32_t bits;
  uint32_t sh;
  uint64_t mask;
};

__host__ __device__ inline uint32_t hash32(uint64_t x, uint32_t key) {
  x = (x ^ key) & 0xffffffffull;
  x = ((x ^ (x >> 16)) * 0x045D9F3Bull) & 0xffffffffull;
  x = ((x ^ (x >> 15)) * 0x0293A5D1ull) & 0xffffffffull;
  return (uint32_t)(x ^ (x >> 16));
}

static PermKeys perm_keys(uint64_t seed, uint32_t salt, uint32_t bits) {
  PermKeys p;
  uint32_t s = (uint32_t)seed * 0x9E3779B1u + salt;
  for (int r = 0; r < 4; ++r) {
    s = s * 1664525u + 1013904223u;
    p.key[r] = s;
  }
  p.bits = bits;
  p.sh = bits / 2 > 1 ? bits / 2 : 1;
  p.mask = (bits >= 64) ? ~0ull : ((1ull << bits) - 1);
  return p;
}

__device__ __forceinline__ uint64_t permute_bits(uint64_t x, const PermKeys& p) {
  const uint64_t mul[4] = {0x2545F491ull, 0x4C957F2Dull, 0x1B873593ull, 0x68E31DA5ull};
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    x = (x + (p.key[r] & p.mask)) & p.mask;
    x = (x * mul[r]) & p.mask;
    x = x ^ (x >> p.sh);
  }
  return x;
}

__global__ __launch_bounds__(256) void rz_gen_removals_kernel(int32_t* __restrict__ R, uint64_t n, uint64_t k,
                                                               PermKeys p) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < k; i += gridDim.x * 256ull) {
    uint64_t y = permute_bits(i, p);
    while (y >= n) y = permute_bits(y, p);  // cycle walking: a bijection of [0, n)
    R[i] = (int32_t)y;
  }
}

__global__ __launch_bounds__(256) void rz_gen_payload_kernel(uint32_t* __restrict__ A, uint64_t n, uint32_t words,
                                                              PermKeys p, uint32_t h1, uint32_t h2, uint32_t h3) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) {
    const uint32_t lo = (uint32_t)permute_bits(i, p);
    uint32_t* out = A + i * words;
    out[0] = lo;
    if (words > 1) out[1] = hash32(i, h1);
    if (words > 2) out[2] = hash32(i, h2);
    if (words > 3) out[3] = hash32(i, h3);
  }
}

static dim3 gen_grid(int cu, uint64_t count) {
  uint64_t g = (count + 255) / 256;
  const uint64_t cap = (uint64_t)cu * 16;
  return dim3((uint32_t)(g < 1 ? 1 : (g > cap ? cap : g)));
}

}  // namespace rz

using namespace rz;

extern "C" {

rz_status rz_gen_uniform_removals(rz_ctx* ctx, int32_t* d_removals, int64_t n, int64_t k, uint64_t seed,
                                  void* stream) {
  if (!ctx || n < 1 || k < 0 || k > n || n >= (1ll << 31) || (k > 0 && !d_removals)) return RZ_EINVAL;
  if (k == 0) return RZ_OK;
  uint32_t bits = 1;
  while (bits < 63 && (1ull << bits) < (uint64_t)n) ++bits;
  const PermKeys p = perm_keys(seed, 1u, bits);
  hipLaunchKernelGGL(rz_gen_removals_kernel, gen_grid(ctx->cu_count, (uint64_t)k), dim3(256), 0,
                     (hipStream_t)stream, d_removals, (uint64_t)n, (uint64_t)k, p);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

rz_status rz_gen_payload(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, uint64_t seed, void* stream) {
  if (!ctx || n < 0 || esz_index(elem_bytes) < 0 || (n > 0 && !d_data)) return RZ_EINVAL;
  if (n == 0) return RZ_OK;
  const uint32_t words = elem_bytes / 4;
  const uint32_t salt = words == 1 ? 3u : words == 2 ? 4u : 5u;
  const PermKeys p = perm_keys(seed, salt, 32);
  const uint32_t s32 = (uint32_t)seed;
  const uint32_t h1 = words == 2 ? s32 * 13u + 1u : s32 * 17u + 2u;
  const uint32_t h2 = s32 * 19u + 9u;
  const uint32_t h3 = s32 * 23u + 4u;
  hipLaunchKernelGGL(rz_gen_payload_kernel, gen_grid(ctx->cu_count, (uint64_t)n), dim3(256), 0, (hipStream_t)stream,
                     (uint32_t*)d_data, (uint64_t)n, words, p, h1, h2, h3);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

}  // extern "C"
