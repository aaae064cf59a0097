// rz_api.hip -- C ABI entry points, context and scratch management.
#include <string.h>

#include <algorithm>
#include <vector>

#include "rz_ctx.hpp"
#include "rz_device.hpp"

namespace rz {

__global__ __launch_bounds__(256) void rz_fill_u32(uint32_t* __restrict__ p, uint32_t v, uint64_t count) {
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < count; i += gridDim.x * 256ull) p[i] = v;
}

rz_status fill_u32(uint32_t* p, uint32_t v, size_t count, hipStream_t s) {
  if (count == 0) return RZ_OK;
  const size_t g = (count + 255) / 256;
  hipLaunchKernelGGL(rz_fill_u32, dim3((uint32_t)(g < 4096 ? g : 4096)), dim3(256), 0, s, p, v, (uint64_t)count);
  RZ_HIP(hipGetLastError());
  return RZ_OK;
}

// Grow a scratch buffer.  New memory is zeroed synchronously; the kernels keep
// it zero.  Growing frees the old buffer, so it must not happen under stream
// capture.  A buffer of the context may have been used on any stream (the
// mixed path runs calls on two), so the whole device is drained before the
// old buffer is freed, and the buffer is detached from the context before
// hipFree is called: a failed free can never leave a pointer that a later
// call would free again.
rz_status ensure(rz_ctx* ctx, DeviceBuffer& b, size_t bytes, hipStream_t s) {
  (void)ctx;
  (void)s;
  if (bytes <= b.bytes) return RZ_OK;
  size_t cap = b.bytes ? b.bytes : 4096;
  while (cap < bytes) cap *= 2;
  RZ_HIP(hipDeviceSynchronize());
  void* old = b.ptr;
  b.ptr = nullptr;
  b.bytes = 0;
  if (old) RZ_HIP(hipFree(old));
  void* p = nullptr;
  if (hipMalloc(&p, cap) != hipSuccess) return RZ_ENOMEM;
  if (hipMemset(p, 0, cap) != hipSuccess) {
    (void)hipFree(p);
    return RZ_EHIP;
  }
  b.ptr = p;
  b.bytes = cap;
  return RZ_OK;
}

rz_status reset_scratch(rz_ctx* ctx) {
  RZ_HIP(hipDeviceSynchronize());
  DeviceBuffer* bufs[] = {&ctx->tail_bits, &ctx->list_bits,  &ctx->counts,    &ctx->orphans,    &ctx->status,
                          &ctx->queue,     &ctx->mstatus,    &ctx->slots,     &ctx->dense_cnt,  &ctx->dense_small,
                          &ctx->dense_ent, &ctx->dense_fine, &ctx->dense_src, &ctx->gstatus,    &ctx->ploaded,
                          &ctx->pgcnt,     &ctx->pgdone,     &ctx->pcur};
  for (DeviceBuffer* b : bufs)
    if (b->ptr) RZ_HIP(hipMemset(b->ptr, 0, b->bytes));
  // every state word (epochs, counters, ready words, the error word); the
  // test hooks behind them stay as the host set them
  RZ_HIP(hipMemset(ctx->ctrl.ptr, 0, (size_t)kCtrlSpinCap * 4));
  rz_status st;
  if (ctx->last_batch && (st = batch_reset_scratch(ctx->last_batch)) != RZ_OK) return st;
  for (rz_batch* b : ctx->batches)
    if ((st = batch_reset_scratch(b)) != RZ_OK) return st;
  RZ_HIP(hipDeviceSynchronize());
  return RZ_OK;
}

rz_status read_errors(rz_ctx* ctx, hipStream_t s, bool all, uint32_t* error_flags) {
  if (all)
    RZ_HIP(hipMemcpyAsync(ctx->ctrl_host, ctx->ctrl.ptr, kCtrlWords * 4, hipMemcpyDeviceToHost, s));
  else
    RZ_HIP(hipMemcpyAsync(ctx->ctrl_host + kCtrlError, (uint32_t*)ctx->ctrl.ptr + kCtrlError, 4,
                          hipMemcpyDeviceToHost, s));
  RZ_HIP(hipStreamSynchronize(s));
  const uint32_t e = ctx->ctrl_host[kCtrlError];
  if (error_flags) *error_flags = e;
  if (!e) return RZ_OK;
  const rz_status st = reset_scratch(ctx);
  return st != RZ_OK ? st : RZ_EDEVICE;
}

rz_status finish_call(rz_ctx* ctx, uint32_t flags, hipStream_t s) {
  if (!(flags & (RZ_STATS | RZ_VALIDATE))) return RZ_OK;
  return read_errors(ctx, s, (flags & RZ_STATS) != 0, nullptr);
}

rz_status prof_mark(rz_ctx* ctx, uint32_t flags, int i, hipStream_t s) {
  if (!(flags & RZ_PROFILE)) return RZ_OK;
  if (!ctx->ev[i]) RZ_HIP(hipEventCreate(&ctx->ev[i]));
  RZ_HIP(hipEventRecord(ctx->ev[i], s));
  if (i == 3) ctx->ev_recorded = true;
  return RZ_OK;
}

rz_tile_config default_config(uint32_t esz, int64_t n, int64_t k) {
  (void)esz;
  (void)n;
  rz_tile_config c;
  c.wg_size = 256;
  c.items = k >= (1 << 22) ? 8 : 4;
  c.max_slices = 2048;
  c.variant = 0;
  c.slice_steps = 4;
  c.chunk = 1;
  return c;
}

static rz_status check_args(rz_ctx* ctx, const void* data, int64_t n, uint32_t esz, const int32_t* R, int64_t k) {
  if (!ctx) return RZ_EINVAL;
  if (n < 0 || k < 0 || k > n || n >= (1ll << 31)) return RZ_EINVAL;
  if (esz_index(esz) < 0) return RZ_EINVAL;
  if (k > 0 && (!data || !R)) return RZ_EINVAL;
  if (data && ((uintptr_t)data % (esz > 8 ? 16 : esz)) != 0) return RZ_EINVAL;
  if (R && ((uintptr_t)R % 4) != 0) return RZ_EINVAL;
  return RZ_OK;
}

}  // namespace rz

using namespace rz;

extern "C" {

rz_status rz_ctx_create(rz_ctx** out, int32_t device) {
  if (!out) return RZ_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return RZ_EHIP;
  DeviceGuard g(device);
  hipDeviceProp_t prop;
  RZ_HIP(hipGetDeviceProperties(&prop, device));
  rz_ctx* ctx = new rz_ctx();
  ctx->device = device;
  ctx->cu_count = prop.multiProcessorCount;
  ctx->lds_per_block = prop.sharedMemPerBlock < 65536 ? (int)prop.sharedMemPerBlock : 65536;
  ctx->lds_max = (int)prop.sharedMemPerBlock;
  strncpy(ctx->arch, prop.gcnArchName, sizeof(ctx->arch) - 1);
  void* p = nullptr;
  if (hipMalloc(&p, kCtrlWords * 4) != hipSuccess) { delete ctx; return RZ_ENOMEM; }
  if (hipMemset(p, 0, kCtrlWords * 4) != hipSuccess) { (void)hipFree(p); delete ctx; return RZ_EHIP; }
  ctx->ctrl.ptr = p;
  ctx->ctrl.bytes = kCtrlWords * 4;
  if (hipHostMalloc((void**)&ctx->ctrl_host, kCtrlWords * 4, hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(p);
    delete ctx;
    return RZ_ENOMEM;
  }
  memset(ctx->ctrl_host, 0, kCtrlWords * 4);
  for (int e = 0; e < 3; ++e)
    for (int b = 0; b < 33; ++b) ctx->table[e][b] = default_config(4u << e, 0, b ? (1ll << b) : 0);
  ctx->batch_cfg = default_config(4, 0, 0);
  ctx->batch_cfg.items = 8;
  ctx->batch_cfg.slice_steps = 2;
  *out = ctx;
  return RZ_OK;
}

rz_status rz_ctx_destroy(rz_ctx* ctx) {
  if (!ctx) return RZ_OK;
  DeviceGuard g(ctx->device);
  (void)hipDeviceSynchronize();
  // batches the caller has not destroyed no longer launch (RZ_EINVAL) but can still be destroyed
  for (rz_batch* b : ctx->batches) batch_detach(b);
  ctx->batches.clear();
  if (ctx->last_batch) rz_batch_destroy(ctx->last_batch);
  DeviceBuffer* bufs[] = {&ctx->tail_bits, &ctx->list_bits, &ctx->counts, &ctx->orphans, &ctx->stamps,
                          &ctx->status,    &ctx->ctrl,      &ctx->queue,  &ctx->mstatus, &ctx->slots,
                          &ctx->dense_cnt, &ctx->dense_small, &ctx->dense_ent, &ctx->dense_fine,
                          &ctx->dense_src, &ctx->dispatch, &ctx->gstatus, &ctx->ploaded, &ctx->pgcnt,
                          &ctx->pgdone, &ctx->pent, &ctx->pcur, &ctx->choose};
  if (ctx->choose_ev) (void)hipEventSynchronize(ctx->choose_ev);  // its copy back has left the staging
  for (DeviceBuffer* b : bufs)
    if (b->ptr) (void)hipFree(b->ptr);
  if (ctx->ctrl_host) (void)hipHostFree(ctx->ctrl_host);
  if (ctx->auto_host) (void)hipHostFree(ctx->auto_host);
  if (ctx->choose_host) (void)hipHostFree(ctx->choose_host);
  if (ctx->choose_ev) (void)hipEventDestroy(ctx->choose_ev);
  for (hipEvent_t e : ctx->ev)
    if (e) (void)hipEventDestroy(e);
  delete ctx;
  return RZ_OK;
}

const char* rz_status_string(rz_status s) {
  switch (s) {
    case RZ_OK: return "ok";
    case RZ_EINVAL: return "invalid argument";
    case RZ_ERANGE: return "removal index out of range";
    case RZ_EDUP: return "duplicate removal index";
    case RZ_ECOOP: return "cooperative launch refused";
    case RZ_EHIP: return "HIP runtime error";
    case RZ_ENOMEM: return "out of device memory";
    case RZ_ENOTUNED: return "no model loaded";
    case RZ_EDEVICE: return "a check on the device failed (output unspecified; scratch reset)";
    case RZ_ETOOBIG: return "k above the fixed scratch of the allocation-free entry point";
  }
  return "unknown status";
}

rz_status rz_device_info(rz_ctx* ctx, int32_t* cu_count, char* arch, int32_t arch_len) {
  if (!ctx) return RZ_EINVAL;
  if (cu_count) *cu_count = ctx->cu_count;
  if (arch && arch_len > 0) {
    strncpy(arch, ctx->arch, (size_t)arch_len - 1);
    arch[arch_len - 1] = 0;
  }
  return RZ_OK;
}

rz_status rz_remove_config(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals,
                           int64_t k, uint32_t flags, const rz_tile_config* cfg, void* stream) {
  rz_status st = check_args(ctx, d_data, n, elem_bytes, d_removals, k);
  if (st != RZ_OK) return st;
  if (!cfg) return RZ_EINVAL;
  if (k == 0 || k == n) return RZ_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  DeviceGuard g(ctx->device);
  hipStream_t s = (hipStream_t)stream;
  if (flags & RZ_VALIDATE) {
    if ((st = launch_validate(ctx, d_removals, n, k, flags, s)) != RZ_OK) return st;
  }
  st = launch_plain(ctx, d_data, n, elem_bytes, d_removals, k, flags, *cfg, s);
  if (st != RZ_OK) return st;
  return finish_call(ctx, flags, s);
}

rz_status rz_remove(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals, int64_t k,
                    uint32_t flags, void* stream) {
  rz_status st = check_args(ctx, d_data, n, elem_bytes, d_removals, k);
  if (st != RZ_OK) return st;
  if (k == 0 || k == n) return RZ_OK;
  const rz_tile_config cfg = ctx->table[esz_index(elem_bytes)][log2_floor(k)];
  return rz_remove_config(ctx, d_data, n, elem_bytes, d_removals, k, flags, &cfg, stream);
}

rz_status rz_remove_sized(rz_ctx* ctx, void* d_data, int64_t n_cap, uint32_t elem_bytes, int32_t* d_removals,
                          int64_t k_cap, const int32_t* d_n, const int32_t* d_k, uint32_t flags, void* stream) {
  rz_status st = check_args(ctx, d_data, n_cap, elem_bytes, d_removals, k_cap);
  if (st != RZ_OK) return st;
  if (!d_n || !d_k || ((uintptr_t)d_n % 4) != 0 || ((uintptr_t)d_k % 4) != 0) return RZ_EINVAL;
  const uint32_t allowed = RZ_PRESERVE_REMOVALS | RZ_SORTED | RZ_ANY_PAIRING | RZ_STATS | RZ_PROFILE;
  if (flags & ~allowed) return RZ_EINVAL;  // RZ_VALIDATE needs k on the host; RZ_COOPERATIVE has no sized form
  if (k_cap == 0) return RZ_OK;  // the device k can only be 0 (or break its bound, which nothing then reports)
  std::lock_guard<std::mutex> lock(ctx->mu);
  DeviceGuard g(ctx->device);
  rz_tile_config cfg;
  if (!sized_shape(ctx, elem_bytes, k_cap, &cfg)) return RZ_EINVAL;
  hipStream_t s = (hipStream_t)stream;
  st = launch_sized(ctx, d_data, n_cap, elem_bytes, d_removals, k_cap, d_n, d_k, flags, cfg, s);
  if (st != RZ_OK) return st;
  return finish_call(ctx, flags, s);
}

rz_status rz_remove_premarked(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals,
                              int64_t k, uint32_t* d_tail_bits, uint32_t flags, const rz_tile_config* cfg,
                              void* stream) {
  rz_status st = check_args(ctx, d_data, n, elem_bytes, d_removals, k);
  if (st != RZ_OK) return st;
  if (k > 0 && (!d_tail_bits || ((uintptr_t)d_tail_bits % 4) != 0)) return RZ_EINVAL;
  if (k == 0) return RZ_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  DeviceGuard g(ctx->device);
  hipStream_t s = (hipStream_t)stream;
  if (k == n) return fill_u32(d_tail_bits, 0u, (size_t)((k + 31) / 32), s);  // nothing moves; leave the bitmap zero
  if (flags & RZ_VALIDATE) {
    if ((st = launch_validate(ctx, d_removals, n, k, flags, s)) != RZ_OK) return st;
  }
  const rz_tile_config c = cfg ? *cfg : ctx->table[esz_index(elem_bytes)][log2_floor(k)];
  st = launch_plain(ctx, d_data, n, elem_bytes, d_removals, k, (flags & ~(uint32_t)RZ_SORTED) | kFlagPremarked, c, s,
                    d_tail_bits);
  if (st != RZ_OK) return st;
  return finish_call(ctx, flags, s);
}

rz_status rz_mark_removals(rz_ctx* ctx, const int32_t* d_removals, int64_t n, int64_t k, uint32_t* d_tail_bits,
                           void* stream) {
  if (!ctx || n < 0 || k < 0 || k > n || n >= (1ll << 31)) return RZ_EINVAL;
  if (k > 0 && (!d_removals || !d_tail_bits)) return RZ_EINVAL;
  DeviceGuard g(ctx->device);
  return launch_mark_removals(ctx, d_removals, n, k, d_tail_bits, (hipStream_t)stream);
}

rz_status rz_remove_u32(rz_ctx* ctx, uint32_t* d_data, int64_t n, int32_t* d_removals, int64_t k, uint32_t flags,
                        void* stream) {
  return rz_remove(ctx, d_data, n, 4, d_removals, k, flags, stream);
}
rz_status rz_remove_u64(rz_ctx* ctx, uint64_t* d_data, int64_t n, int32_t* d_removals, int64_t k, uint32_t flags,
                        void* stream) {
  return rz_remove(ctx, d_data, n, 8, d_removals, k, flags, stream);
}
rz_status rz_remove_rec16(rz_ctx* ctx, void* d_data, int64_t n, int32_t* d_removals, int64_t k, uint32_t flags,
                          void* stream) {
  return rz_remove(ctx, d_data, n, 16, d_removals, k, flags, stream);
}

rz_status rz_last_stats(rz_ctx* ctx, rz_stats* out) {
  if (!ctx || !out) return RZ_EINVAL;
  memset(out, 0, sizeof(*out));
  const uint32_t* h = ctx->ctrl_host;
  out->tail_removals = (int64_t)(h[kCtrlStatTail] | ((uint64_t)h[kCtrlStatTail + 1] << 32));
  out->orphan_pairs = (int64_t)(h[kCtrlStatOrphans] | ((uint64_t)h[kCtrlStatOrphans + 1] << 32));
  out->error_flags = (int32_t)h[kCtrlError];
  return RZ_OK;
}

rz_status rz_remove_batched(rz_ctx* ctx, const rz_list_desc* lists, int32_t nlists, uint32_t flags, void* stream) {
  if (!ctx) return RZ_EINVAL;
  if (nlists < 0 || (nlists > 0 && !lists) || nlists >= (1 << 28)) return RZ_EINVAL;
  std::lock_guard<std::mutex> lock(ctx->mu);
  DeviceGuard g(ctx->device);
  return batch_plan_and_launch(ctx, lists, nlists, flags, (hipStream_t)stream);
}

rz_status rz_last_profile(rz_ctx* ctx, float ms[4]) {
  if (!ctx || !ms) return RZ_EINVAL;
  if (!ctx->ev_recorded) return RZ_EINVAL;
  DeviceGuard g(ctx->device);
  RZ_HIP(hipEventSynchronize(ctx->ev[3]));
  RZ_HIP(hipEventElapsedTime(&ms[0], ctx->ev[0], ctx->ev[1]));
  RZ_HIP(hipEventElapsedTime(&ms[1], ctx->ev[1], ctx->ev[2]));
  RZ_HIP(hipEventElapsedTime(&ms[2], ctx->ev[2], ctx->ev[3]));
  RZ_HIP(hipEventElapsedTime(&ms[3], ctx->ev[0], ctx->ev[3]));
  return RZ_OK;
}

rz_status rz_last_stages(rz_ctx* ctx, float out[16]) {
  if (!ctx || !out) return RZ_EINVAL;
  memset(out, 0, 16 * sizeof(float));
  if (!ctx->stamp_groups || !ctx->stamps.ptr) return RZ_EINVAL;
  DeviceGuard g(ctx->device);
  const uint32_t groups = ctx->stamp_groups, nm = ctx->stamp_marks;
  std::vector<uint64_t> h((size_t)groups * kStampsPerGroup);
  RZ_HIP(hipDeviceSynchronize());
  RZ_HIP(hipMemcpy(h.data(), ctx->stamps.ptr, h.size() * 8, hipMemcpyDeviceToHost));
  int khz = 0;
  RZ_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
  if (khz <= 0) return RZ_EHIP;
  const double us_per_tick = 1e3 / (double)khz;
  auto row = [&](uint32_t gidx, int i) { return h[(size_t)gidx * kStampsPerGroup + i]; };
  uint64_t t0 = ~0ull;
  for (uint32_t gi = 0; gi < groups; ++gi)
    if (row(gi, kStampStart) && row(gi, kStampStart) < t0) t0 = row(gi, kStampStart);
  if (t0 == ~0ull) return RZ_EINVAL;
  auto rel = [&](uint64_t t) { return t ? (float)((double)(t - t0) * us_per_tick) : 0.f; };
  uint64_t marked = 0, raised = 0;
  for (uint32_t gi = 0; gi < nm; ++gi) {
    marked = std::max(marked, row(gi, kStampMarked));
    raised = std::max(raised, row(gi, kStampRaised));
  }
  uint64_t first = ~0ull, last = 0, mx[6] = {0, 0, 0, 0, 0, 0};
  double sum[6] = {0, 0, 0, 0, 0, 0};
  const uint32_t ns = groups - nm;
  for (uint32_t gi = nm; gi < groups; ++gi) {
    const uint64_t st0 = row(gi, kStampStart);
    first = std::min(first, st0);
    last = std::max(last, st0);
    uint64_t prev = st0;
    for (int i = kStampReady; i <= kStampExit; ++i) {
      const uint64_t t = row(gi, i);
      if (!t) continue;  // no ready wait in the two-kernel form
      mx[i] = std::max(mx[i], t);
      sum[i] += (double)(t - prev) * us_per_tick;
      prev = t;
    }
  }
  out[0] = rel(marked);
  out[1] = rel(raised);
  out[2] = rel(first);
  out[3] = rel(last);
  for (int i = kStampReady; i <= kStampExit; ++i) {
    out[3 + i] = rel(mx[i]);                            // out[4..8]: last slice to reach the stage
    out[8 + i] = ns ? (float)(sum[i] / (double)ns) : 0.f;  // out[9..13]: mean time a slice spends in it
  }
  out[14] = (float)nm;
  out[15] = (float)ns;
  return RZ_OK;
}

rz_status rz_partition_on(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, const int32_t* d_removals,
                          int64_t k, uint32_t flags, void* stream) {
  rz_status st = check_args(ctx, d_data, n, elem_bytes, d_removals, k);
  if (st != RZ_OK) return st;
  if (k == 0 || n == 0) return RZ_OK;
  std::lock_guard<std::mutex> lock(ctx->mu);
  DeviceGuard g(ctx->device);
  hipStream_t s = (hipStream_t)stream;
  if (flags & RZ_VALIDATE) {
    if ((st = launch_validate(ctx, d_removals, n, k, flags, s)) != RZ_OK) return st;
  }
  if ((st = launch_partition(ctx, d_data, n, elem_bytes, d_removals, k, flags, s)) != RZ_OK) return st;
  return finish_call(ctx, flags, s);
}

rz_status rz_check_errors(rz_ctx* ctx, void* stream, uint32_t* error_flags) {
  if (!ctx) return RZ_EINVAL;
  std::lock_guard<std::mutex> lock(ctx->mu);
  DeviceGuard g(ctx->device);
  return read_errors(ctx, (hipStream_t)stream, false, error_flags);
}

rz_status rz_debug_set_waits(rz_ctx* ctx, uint32_t spin_cap, uint32_t stall) {
  if (!ctx) return RZ_EINVAL;
  std::lock_guard<std::mutex> lock(ctx->mu);
  DeviceGuard g(ctx->device);
  const uint32_t w[2] = {spin_cap, stall};
  RZ_HIP(hipDeviceSynchronize());
  RZ_HIP(hipMemcpy((uint32_t*)ctx->ctrl.ptr + kCtrlSpinCap, w, sizeof(w), hipMemcpyHostToDevice));
  return RZ_OK;
}

}  // extern "C"
