// rz_tune.hip -- autotuner of launch shapes.
//
// For one (element size, n, k) the tuner builds a synthetic list and removal
// list on the device (rz_gen_*), times every candidate launch shape with
// events on a private stream (restoring the inputs and, for cold timing,
// reading a 1 GiB scrub buffer before each call), and installs the fastest in
// the context's table under (element size, floor(log2 k)).  The batch tuner
// does the same for the persistent queue on a synthetic batch of lists.  The
// table is saved and loaded as plain text lines.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <vector>

#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

__global__ __launch_bounds__(256) void rz_scrub_kernel(const uint4* __restrict__ p, uint64_t n16,
                                                        uint32_t* __restrict__ sink) {
  uint32_t acc = 0;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n16; i += gridDim.x * 256ull) {
    const uint4 v = p[i];
    acc ^= v.x ^ v.y ^ v.z ^ v.w;
  }
  acc = wave_sum(acc);
  if (acc == 0x9e3779b9u && lane_id() == 0) sink[0] = acc;  // keeps the loads alive; practically never true
}

// Counts the 4-byte words that differ between a and b inside the ranges
// (word offset, word count) into *mismatches.
__global__ __launch_bounds__(256) void rz_compare_ranges(const uint32_t* __restrict__ a,
                                                          const uint32_t* __restrict__ b,
                                                          const uint64_t* __restrict__ ranges, uint32_t nranges,
                                                          uint32_t* __restrict__ mismatches) {
  uint32_t bad = 0;
  for (uint32_t r = blockIdx.x; r < nranges; r += gridDim.x) {
    const uint64_t off = ranges[2 * r], cnt = ranges[2 * r + 1];
    for (uint64_t i = threadIdx.x; i < cnt; i += 256) bad += a[off + i] != b[off + i] ? 1u : 0u;
  }
  bad = wave_sum(bad);
  if (lane_id() == 0 && bad) atomicAdd(mismatches, bad);
}

struct Scratch {
  std::vector<void*> ptrs;
  ~Scratch() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  void* alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return p;
  }
};

// Whether a candidate's last call left the same survivors as the reference
// result and no device error: the word ranges are compared on the device, and
// the context's error word is read (and, if set, cleared with the scratch).
static rz_status candidate_ok(rz_ctx* ctx, const void* got, const void* ref, const uint64_t* d_ranges,
                              uint32_t nranges, uint32_t* d_count, hipStream_t s, bool* ok) {
  *ok = false;
  RZ_HIP(hipMemsetAsync(d_count, 0, 4, s));
  hipLaunchKernelGGL(rz_compare_ranges, dim3(std::min<uint32_t>(nranges, 4096u)), dim3(256), 0, s,
                     (const uint32_t*)got, (const uint32_t*)ref, d_ranges, nranges, d_count);
  RZ_HIP(hipGetLastError());
  uint32_t bad = 0;
  RZ_HIP(hipMemcpyAsync(&bad, d_count, 4, hipMemcpyDeviceToHost, s));
  RZ_HIP(hipStreamSynchronize(s));
  uint32_t err = 0;
  const rz_status st = rz_check_errors(ctx, s, &err);
  if (st != RZ_OK && st != RZ_EDEVICE) return st;
  *ok = bad == 0 && err == 0;
  return RZ_OK;
}

static float median(std::vector<float> v) {
  std::sort(v.begin(), v.end());
  return v.empty() ? 0.f : v[v.size() / 2];
}

static const rz_tile_config kCandidates[] = {
    // wg, items, max_slices, variant, slice_steps, chunk
    {256, 4, 512, 0, 0, 1},   {512, 8, 4096, 0, 0, 1},  {256, 8, 4096, 0, 0, 1},                            // three kernels
    {256, 4, 4096, 3, 0, 1},  {256, 8, 4096, 3, 0, 1},  {512, 8, 4096, 3, 0, 1},                            // ordered
    {256, 8, 0, 2, 1, 1},                                                                                   // queue
    {256, 4, 512, 4, 0, 1},   {256, 4, 2048, 4, 0, 1},  {256, 8, 512, 4, 0, 1},  {512, 8, 4096, 4, 0, 1},   // fused
    {256, 2, 4096, 4, 0, 1},  {256, 8, 4096, 4, 0, 1},
    {256, 4, 512, 6, 10, 1},  {256, 2, 4096, 6, 10, 1}, {256, 4, 512, 6, 12, 1}, {256, 8, 512, 6, 14, 1},  // one launch
    {512, 4, 512, 6, 10, 1},  {512, 4, 512, 6, 12, 1},  {256, 4, 1024, 6, 11, 1},
    {1024, 8, 1024, 5, 11, 2}, {1024, 8, 1024, 5, 12, 2}, {1024, 4, 1024, 5, 11, 2}, {512, 8, 2048, 5, 11, 2},  // dense
    {1024, 8, 1024, 5, 11, 0},
};

static const rz_tile_config kBatchCandidates[] = {
    {256, 4, 0, 2, 1, 1}, {256, 4, 0, 2, 4, 1}, {256, 8, 0, 2, 1, 1}, {256, 8, 0, 2, 2, 1},
    {256, 8, 0, 2, 4, 1}, {512, 8, 0, 2, 1, 1}, {512, 8, 0, 2, 2, 1}, {512, 4, 0, 2, 2, 1},
    {256, 8, 0, 2, 2, 4}, {512, 8, 0, 2, 1, 2},
    // fused items (a list's slices pair their orphans through look-back and exchange words)
    {256, 4, 0, 4, 4, 1}, {256, 8, 0, 4, 1, 1}, {256, 8, 0, 4, 2, 1}, {512, 8, 0, 4, 1, 1},
    {512, 8, 0, 4, 2, 1}, {512, 4, 0, 4, 2, 1}, {512, 8, 0, 4, 1, 2},
    // bucketed tail marks for lists of 2^17 removals and more
    {512, 8, 0, 5, 1, 1}, {512, 8, 0, 5, 2, 1}, {256, 8, 0, 5, 1, 1}, {512, 4, 0, 5, 2, 1},
};

}  // namespace rz

using namespace rz;

extern "C" {

rz_status rz_tune_candidates(int32_t batch, rz_tile_config* out, int32_t max, int32_t* count) {
  const rz_tile_config* src = batch ? kBatchCandidates : kCandidates;
  const int32_t n = batch ? (int32_t)(sizeof(kBatchCandidates) / sizeof(kBatchCandidates[0]))
                          : (int32_t)(sizeof(kCandidates) / sizeof(kCandidates[0]));
  if (count) *count = n;
  for (int32_t i = 0; out && i < n && i < max; ++i) out[i] = src[i];
  return RZ_OK;
}

rz_status rz_table_get(rz_ctx* ctx, uint32_t elem_bytes, int32_t log2k, rz_tile_config* out) {
  if (!ctx || !out || esz_index(elem_bytes) < 0 || log2k < 0 || log2k > 32) return RZ_EINVAL;
  *out = ctx->table[esz_index(elem_bytes)][log2k];
  return RZ_OK;
}

rz_status rz_table_set(rz_ctx* ctx, uint32_t elem_bytes, int32_t log2k, const rz_tile_config* cfg) {
  if (!ctx || !cfg || esz_index(elem_bytes) < 0 || log2k < 0 || log2k > 32) return RZ_EINVAL;
  ctx->table[esz_index(elem_bytes)][log2k] = *cfg;
  return RZ_OK;
}

rz_status rz_batch_config_get(rz_ctx* ctx, rz_tile_config* out) {
  if (!ctx || !out) return RZ_EINVAL;
  *out = ctx->batch_cfg;
  return RZ_OK;
}

rz_status rz_batch_config_set(rz_ctx* ctx, const rz_tile_config* cfg) {
  if (!ctx || !cfg || (cfg->variant != 2 && cfg->variant != 4 && cfg->variant != 5)) return RZ_EINVAL;
  ctx->batch_cfg = *cfg;
  return RZ_OK;
}

// Text format, one line per entry:
//   plain <elem_bytes> <log2k> <wg> <items> <max_slices> <variant> <slice_steps> <chunk>
//   batch <wg> <items> <max_slices> <variant> <slice_steps> <chunk>
rz_status rz_table_save(rz_ctx* ctx, const char* path) {
  if (!ctx || !path) return RZ_EINVAL;
  FILE* f = fopen(path, "w");
  if (!f) return RZ_EINVAL;
  fprintf(f, "# redzone_amd launch-shape table (%s, %d CUs)\n", ctx->arch, ctx->cu_count);
  for (int e = 0; e < 3; ++e)
    for (int b = 0; b <= 32; ++b) {
      const rz_tile_config& c = ctx->table[e][b];
      fprintf(f, "plain %d %d %d %d %d %d %d %d\n", 4 << e, b, c.wg_size, c.items, c.max_slices, c.variant,
              c.slice_steps, c.chunk);
    }
  const rz_tile_config& c = ctx->batch_cfg;
  fprintf(f, "batch %d %d %d %d %d %d\n", c.wg_size, c.items, c.max_slices, c.variant, c.slice_steps, c.chunk);
  fclose(f);
  return RZ_OK;
}

rz_status rz_table_load(rz_ctx* ctx, const char* path) {
  if (!ctx || !path) return RZ_EINVAL;
  FILE* f = fopen(path, "r");
  if (!f) return RZ_EINVAL;
  char line[256];
  rz_status st = RZ_OK;
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    rz_tile_config c;
    int e = 0, b = 0;
    if (sscanf(line, "plain %d %d %d %d %d %d %d %d", &e, &b, &c.wg_size, &c.items, &c.max_slices, &c.variant,
               &c.slice_steps, &c.chunk) == 8) {
      if (esz_index((uint32_t)e) < 0 || b < 0 || b > 32) { st = RZ_EINVAL; break; }
      ctx->table[esz_index((uint32_t)e)][b] = c;
    } else if (sscanf(line, "batch %d %d %d %d %d %d", &c.wg_size, &c.items, &c.max_slices, &c.variant,
                      &c.slice_steps, &c.chunk) == 6) {
      ctx->batch_cfg = c;
    } else {
      st = RZ_EINVAL;
      break;
    }
  }
  fclose(f);
  return st;
}

// Times each candidate (median of `reps` calls after one warm-up call) on a
// synthetic list; all_ms (ncand entries, may be null) receives the medians.
// A candidate whose result differs from the three-kernel form's, or that sets
// a device error bit, is rejected (-2).  The fastest of the rest is installed
// in the table.  Allocates its own inputs.
rz_status rz_tune(rz_ctx* ctx, uint32_t elem_bytes, int64_t n, int64_t k, const rz_tile_config* cands,
                  int32_t ncand, int32_t reps, int32_t cold, rz_tile_config* best, float* best_ms, float* all_ms) {
  if (!ctx || esz_index(elem_bytes) < 0 || n < 1 || k < 1 || k >= n || n >= (1ll << 31) || reps < 1) return RZ_EINVAL;
  if (!cands) {
    cands = kCandidates;
    ncand = (int32_t)(sizeof(kCandidates) / sizeof(kCandidates[0]));
  }
  if (ncand < 1) return RZ_EINVAL;
  int dev_prev = -1;
  (void)hipGetDevice(&dev_prev);
  if (dev_prev != ctx->device) (void)hipSetDevice(ctx->device);
  Scratch sc;
  const size_t bytes = (size_t)n * elem_bytes;
  void* data = sc.alloc(bytes);
  void* pristine = sc.alloc(bytes);
  int32_t* R = (int32_t*)sc.alloc((size_t)k * 4);
  int32_t* R0 = (int32_t*)sc.alloc((size_t)k * 4);
  const uint64_t scrub_bytes = cold ? (1ull << 30) : 0;
  uint4* scrub = cold ? (uint4*)sc.alloc(scrub_bytes) : nullptr;
  uint32_t* sink = (uint32_t*)sc.alloc(256);
  void* ref = sc.alloc(bytes);
  uint64_t* d_range = (uint64_t*)sc.alloc(16);
  uint32_t* d_count = (uint32_t*)sc.alloc(256);
  if (!data || !pristine || !R || !R0 || (cold && !scrub) || !sink || !ref || !d_range || !d_count) return RZ_ENOMEM;
  hipStream_t s;
  RZ_HIP(hipStreamCreate(&s));
  hipEvent_t e0, e1;
  RZ_HIP(hipEventCreate(&e0));
  RZ_HIP(hipEventCreate(&e1));
  rz_status st = rz_gen_payload(ctx, pristine, n, elem_bytes, 0x5eed0001ull, s);
  if (st == RZ_OK) st = rz_gen_uniform_removals(ctx, R0, n, k, 0x5eed0002ull, s);
  if (cold && st == RZ_OK) (void)hipMemsetAsync(scrub, 0x5a, scrub_bytes, s);
  // the reference result every candidate must reproduce: the three-kernel form
  // with the default shape
  const uint64_t range[2] = {0, (uint64_t)(n - k) * elem_bytes / 4};
  const rz_tile_config ref_cfg = default_config(elem_bytes, n, k);
  if (st == RZ_OK &&
      (hipMemcpyAsync(ref, pristine, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
       hipMemcpyAsync(R, R0, (size_t)k * 4, hipMemcpyDeviceToDevice, s) != hipSuccess ||
       hipMemcpyAsync(d_range, range, 16, hipMemcpyHostToDevice, s) != hipSuccess))
    st = RZ_EHIP;
  if (st == RZ_OK) st = rz_remove_config(ctx, ref, n, elem_bytes, R, k, 0, &ref_cfg, s);
  if (st == RZ_OK) {
    uint32_t err = 0;
    st = rz_check_errors(ctx, s, &err);
  }
  int best_i = -1;
  float best_t = 0.f;
  for (int32_t c = 0; c < ncand && st == RZ_OK; ++c) {
    std::vector<float> ts;
    for (int32_t r = 0; r <= reps && st == RZ_OK; ++r) {
      if (hipMemcpyAsync(data, pristine, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(R, R0, (size_t)k * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) {
        st = RZ_EHIP;
        break;
      }
      if (cold)
        hipLaunchKernelGGL(rz_scrub_kernel, dim3((uint32_t)ctx->cu_count * 8), dim3(256), 0, s, scrub,
                           scrub_bytes / 16, sink);
      (void)hipEventRecord(e0, s);
      st = rz_remove_config(ctx, data, n, elem_bytes, R, k, 0, &cands[c], s);
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) st = RZ_EHIP;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0) ts.push_back(ms);
    }
    if (st != RZ_OK) {
      // a shape the device refuses (e.g. a cooperative grid too large) is skipped
      if (st == RZ_ECOOP || st == RZ_EINVAL) {
        st = RZ_OK;
        if (all_ms) all_ms[c] = -1.f;
        continue;
      }
      break;
    }
    // the last call's survivors must equal the reference's, with no device error
    bool ok = false;
    if ((st = candidate_ok(ctx, data, ref, d_range, 1, d_count, s, &ok)) != RZ_OK) break;
    if (!ok) {
      if (all_ms) all_ms[c] = -2.f;
      continue;
    }
    const float m = median(ts);
    if (all_ms) all_ms[c] = m;
    if (best_i < 0 || m < best_t) {
      best_i = c;
      best_t = m;
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  if (dev_prev >= 0 && dev_prev != ctx->device) (void)hipSetDevice(dev_prev);
  if (st != RZ_OK) return st;
  if (best_i < 0) return RZ_EINVAL;
  ctx->table[esz_index(elem_bytes)][log2_floor(k)] = cands[best_i];
  if (best) *best = cands[best_i];
  if (best_ms) *best_ms = best_t;
  return RZ_OK;
}

// Batch tuner: nlists lists with n log-uniform in [n_min, n_max] and
// k = floor(r n), r uniform in [0, r_max] (a config-4-like batch), all of
// elem_bytes; times create-once / launch-many of each candidate.
rz_status rz_tune_batch(rz_ctx* ctx, uint32_t elem_bytes, int32_t nlists, int64_t n_min, int64_t n_max,
                        double r_max, int32_t reps, const rz_tile_config* cands, int32_t ncand,
                        rz_tile_config* best, float* best_ms, float* all_ms) {
  if (!ctx || esz_index(elem_bytes) < 0 || nlists < 1 || n_min < 1 || n_max < n_min || n_max >= (1ll << 31) ||
      r_max < 0 || r_max > 1 || reps < 1)
    return RZ_EINVAL;
  if (!cands) {
    cands = kBatchCandidates;
    ncand = (int32_t)(sizeof(kBatchCandidates) / sizeof(kBatchCandidates[0]));
  }
  int dev_prev = -1;
  (void)hipGetDevice(&dev_prev);
  if (dev_prev != ctx->device) (void)hipSetDevice(ctx->device);
  std::mt19937_64 rng(0x7a11ull);
  std::uniform_real_distribution<double> u01(0.0, 1.0);
  std::vector<int64_t> ns(nlists), ks(nlists), off(nlists + 1, 0), roff(nlists + 1, 0);
  const double l0 = log((double)n_min), l1 = log((double)n_max);
  for (int32_t i = 0; i < nlists; ++i) {
    ns[i] = (int64_t)exp(l0 + (l1 - l0) * u01(rng));
    if (ns[i] < 1) ns[i] = 1;
    ks[i] = (int64_t)(u01(rng) * r_max * (double)ns[i]);
    off[i + 1] = off[i] + ns[i];
    roff[i + 1] = roff[i] + ks[i];
  }
  Scratch sc;
  const size_t bytes = (size_t)off[nlists] * elem_bytes;
  char* data = (char*)sc.alloc(bytes);
  char* pristine = (char*)sc.alloc(bytes);
  int32_t* R = (int32_t*)sc.alloc((size_t)roff[nlists] * 4 + 4);
  int32_t* R0 = (int32_t*)sc.alloc((size_t)roff[nlists] * 4 + 4);
  char* ref = (char*)sc.alloc(bytes);
  uint64_t* d_ranges = (uint64_t*)sc.alloc((size_t)nlists * 16);
  uint32_t* d_count = (uint32_t*)sc.alloc(256);
  if (!data || !pristine || !R || !R0 || !ref || !d_ranges || !d_count) return RZ_ENOMEM;
  hipStream_t s;
  RZ_HIP(hipStreamCreate(&s));
  rz_status st = RZ_OK;
  std::vector<rz_list_desc> descs(nlists);
  std::vector<uint64_t> ranges(2 * (size_t)nlists);
  for (int32_t i = 0; i < nlists && st == RZ_OK; ++i) {
    st = rz_gen_payload(ctx, pristine + (size_t)off[i] * elem_bytes, ns[i], elem_bytes, 100 + i, s);
    if (st == RZ_OK) st = rz_gen_uniform_removals(ctx, R0 + roff[i], ns[i], ks[i], 200 + i, s);
    descs[i].data = data + (size_t)off[i] * elem_bytes;
    descs[i].removals = R + roff[i];
    descs[i].n = ns[i];
    descs[i].k = ks[i];
    descs[i].elem_bytes = elem_bytes;
    descs[i].method = 0;
    ranges[2 * i] = (uint64_t)off[i] * elem_bytes / 4;  // the list's survivors, in words
    ranges[2 * i + 1] = (uint64_t)(ns[i] - ks[i]) * elem_bytes / 4;
  }
  // the reference result: every list by the three-kernel form with the default shape
  if (st == RZ_OK &&
      (hipMemcpyAsync(ref, pristine, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
       hipMemcpyAsync(R, R0, (size_t)roff[nlists] * 4, hipMemcpyDeviceToDevice, s) != hipSuccess ||
       hipMemcpyAsync(d_ranges, ranges.data(), (size_t)nlists * 16, hipMemcpyHostToDevice, s) != hipSuccess))
    st = RZ_EHIP;
  for (int32_t i = 0; i < nlists && st == RZ_OK; ++i) {
    const rz_tile_config rc = default_config(elem_bytes, ns[i], ks[i]);
    st = rz_remove_config(ctx, ref + (size_t)off[i] * elem_bytes, ns[i], elem_bytes, R + roff[i], ks[i], 0, &rc, s);
  }
  if (st == RZ_OK) {
    uint32_t err = 0;
    st = rz_check_errors(ctx, s, &err);
  }
  hipEvent_t e0, e1;
  RZ_HIP(hipEventCreate(&e0));
  RZ_HIP(hipEventCreate(&e1));
  int best_i = -1;
  float best_t = 0.f;
  for (int32_t c = 0; c < ncand && st == RZ_OK; ++c) {
    rz_batch* b = nullptr;
    st = rz_batch_create(ctx, descs.data(), nlists, 0, &cands[c], &b);
    if (st != RZ_OK) break;
    std::vector<float> ts;
    for (int32_t r = 0; r <= reps && st == RZ_OK; ++r) {
      if (hipMemcpyAsync(data, pristine, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
          hipMemcpyAsync(R, R0, (size_t)roff[nlists] * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) {
        st = RZ_EHIP;
        break;
      }
      (void)hipEventRecord(e0, s);
      st = rz_batch_launch(b, s);
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) st = RZ_EHIP;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (r > 0) ts.push_back(ms);
    }
    rz_batch_destroy(b);
    if (st != RZ_OK) break;
    bool ok = false;
    if ((st = candidate_ok(ctx, data, ref, d_ranges, (uint32_t)nlists, d_count, s, &ok)) != RZ_OK) break;
    if (!ok) {
      if (all_ms) all_ms[c] = -2.f;
      continue;
    }
    const float m = median(ts);
    if (all_ms) all_ms[c] = m;
    if (best_i < 0 || m < best_t) {
      best_i = c;
      best_t = m;
    }
  }
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  if (dev_prev >= 0 && dev_prev != ctx->device) (void)hipSetDevice(dev_prev);
  if (st != RZ_OK) return st;
  if (best_i < 0) return RZ_EINVAL;
  ctx->batch_cfg = cands[best_i];
  if (best) *best = cands[best_i];
  if (best_ms) *best_ms = best_t;
  return RZ_OK;
}

}  // extern "C"
