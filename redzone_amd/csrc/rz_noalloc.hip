// rz_noalloc.hip -- rz_remove_noalloc: Redzone removal that allocates nothing.
//
// For callers that cannot call hipMalloc (code under graph capture, tight
// serving loops, harnesses that forbid allocation).  Every word of scratch the
// fused form needs -- the control block, the two halves of look-back status
// words, the tail bitmap and the orphan exchange words -- lives in one
// file-scope device array, `scratch`, of fixed size, which each GPU the library
// runs on holds until the process ends.  Its size bounds k at
// RZ_NOALLOC_MAX_K removals.
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// first call on a device (with that device current) asks the runtime for the
// array's device address with hipGetSymbolAddress, keeps it per device, and
// only that address reaches the kernels.
//
// Every call on one GPU uses that GPU's copy of the array, and a call leaves
// it as the kernels leave it: the bitmap words cleared by the slices, the
// exchange words tagged with the call epoch, the look-back half of the next
// call zeroed.  A prologue kernel still clears the bitmap words of the call's
// k, the error word and the exit counters before the call's kernels run, so a
// call never depends on the previous one having finished cleanly.  Calls that
// overlap on one GPU (two streams, two threads) race on the array.
#include <mutex>

#include "rz_ctx.hpp"

namespace rz {

namespace {

constexpr int64_t kMaxK = RZ_NOALLOC_MAX_K;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }
// layout of the array (bytes)
constexpr size_t kOffCtrl = 0;
constexpr size_t kOffStatus = align_up(kOffCtrl + (size_t)kCtrlWords * 4);
constexpr size_t kOffBits = align_up(kOffStatus + 2 * (size_t)kMaxSlices * 8);
constexpr size_t kOffSlots = align_up(kOffBits + (size_t)((kMaxK + 31) / 32) * 4);
constexpr size_t kScratchBytes = align_up(kOffSlots + ((size_t)kMaxK / 2 + 1) * 8);
static_assert(kScratchBytes < (17u << 20), "the fixed scratch stays below 17 MiB");

constexpr int kMaxDevices = 64;

}  // namespace

// Zero-initialised when the code object is loaded on a device.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;
unsigned char* g_ptr[kMaxDevices] = {};

// The device address of this device's copy of `scratch` (looked up once per
// device, with the device current).
rz_status scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return RZ_EINVAL;
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RZ_EHIP;
    g_ptr[dev] = (unsigned char*)p;
  }
  *out = g_ptr[dev];
  return RZ_OK;
}

// Launch shape by k (the tuned table's choice for 4-byte items): the one-launch
// fused form for small k, the mark kernel then the fused move-and-pair kernel
// from 2^19 removals on.
rz_tile_config noalloc_shape(int64_t k) {
  rz_tile_config c;
  c.wg_size = 256;
  c.max_slices = 512;
  c.chunk = 1;
  if (k < (1 << 19)) {
    c.items = 4;
    c.variant = 6;
    c.slice_steps = 12;  // 4,096-removal mark tiles
  } else {
    c.items = 8;
    c.variant = 4;
    c.slice_steps = 0;
  }
  return c;
}

}  // namespace

}  // namespace rz

using namespace rz;

extern "C" {

rz_status rz_remove_noalloc(void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals, int64_t k,
                            uint32_t flags, uint32_t* d_error_flags, void* stream) {
  if (n < 0 || k < 0 || k > n || n >= (1ll << 31) || esz_index(elem_bytes) < 0) return RZ_EINVAL;
  if (flags & ~(uint32_t)(RZ_PRESERVE_REMOVALS | RZ_SORTED | RZ_ANY_PAIRING)) return RZ_EINVAL;
  if (k > kMaxK) return RZ_ETOOBIG;
  if (k > 0 && (!d_data || !d_removals)) return RZ_EINVAL;
  if (d_data && ((uintptr_t)d_data % (elem_bytes > 8 ? 16 : elem_bytes)) != 0) return RZ_EINVAL;
  if ((d_removals && ((uintptr_t)d_removals % 4) != 0) || ((uintptr_t)d_error_flags % 4) != 0) return RZ_EINVAL;
  if (k == 0 || k == n) return RZ_OK;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return RZ_EHIP;
  unsigned char* base = nullptr;
  rz_status st = scratch_address(dev, &base);
  if (st != RZ_OK) return st;
  FusedScratch fs;
  fs.ctrl = (uint32_t*)(base + kOffCtrl);
  fs.mstatus = (uint64_t*)(base + kOffStatus);
  fs.bits = (uint32_t*)(base + kOffBits);
  fs.slots = (uint64_t*)(base + kOffSlots);
  fs.err_out = d_error_flags;
  // the fused forms never write the removal list, so RZ_PRESERVE_REMOVALS holds by itself
  return launch_fused_static(d_data, n, elem_bytes, d_removals, k, flags, noalloc_shape(k), fs, (hipStream_t)stream);
}

}  // extern "C"
