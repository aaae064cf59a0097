// rz_ctx.hpp -- host-side context: device scratch, tuning table, launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "../../include/redzone.h"

namespace rz {

// Words of the control block (uint32 index).  Each group sits on its own
// 128-byte line so that polling one does not contend with updates of another.
enum CtrlWord : int {
  kCtrlEpoch = 0,         // flag epoch of the O(n) partition (bumped by its mark kernel) and of the
                          // one-pass gather (advanced by its last tile)
  kCtrlError = 32,        // device-detected contract violations (bit flags)
  kCtrlBarCount = 64,     // uint64 arrival counter of the in-kernel grid barrier
  kCtrlBarEpoch = 96,     // launch epoch of the in-kernel grid barrier
  kCtrlStatTail = 128,    // uint64: tail removals k_z (RZ_STATS)
  kCtrlStatOrphans = 160, // uint64: orphan pairs (RZ_STATS)
  kCtrlValidate = 192,    // validation flags: 1 = out of range, 2 = duplicate, 4 = not ascending
  kCtrlMoveEpoch = 224,   // call counter of the fused move-and-pair form (bumped by its last workgroup)
  kCtrlMoveExit = 240,    // workgroups of the fused kernel that have finished (reset by the last)
  kCtrlMarkDone = 256,    // groups of mark workgroups of the one-launch form that are done (reset by the last)
  kCtrlMarkReady = 288,   // 8 copies, one line apart (one per XCD): move-pair epoch + 1 once every mark is in
  kCtrlGatherExit = 544,  // tiles of the one-pass gather that have finished (reset by the last)
  kCtrlMarkGroup = 576,   // 8 copies, one line apart: mark workgroups of the one-launch form done, per
                          // group of ids equal mod 8 (reset by the last of the group)
  kCtrlSpinCap = 832,     // test hook: polls before a bounded wait gives up (0 = the wait's own limit)
  kCtrlStall = 833,       // test hook: nonzero makes the Redzone waits see their condition as unmet
  kCtrlWords = 864
};

enum ErrorBits : uint32_t {
  kErrOrphanMismatch = RZ_ERR_ORPHAN_MISMATCH,   // dest and source orphan totals differ (duplicate or bad input)
  kErrBarrierTimeout = RZ_ERR_WAIT_TIMEOUT,      // an in-kernel wait gave up (workgroups not co-resident)
  kErrLookbackTimeout = RZ_ERR_LOOKBACK_TIMEOUT, // a look-back wait gave up
  kErrSizes = RZ_ERR_SIZES,                      // device-resident n, k outside the call's bounds
};

// Polls a bounded wait makes before it gives up: the control block's cap when
// the host set one, else the wait's own limit.
__device__ __forceinline__ uint32_t spin_cap(const uint32_t* ctrl, uint32_t own) {
  const uint32_t c = __hip_atomic_load(&ctrl[kCtrlSpinCap], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return c ? c : own;
}
__device__ __forceinline__ bool stalled(const uint32_t* ctrl) {
  return __hip_atomic_load(&ctrl[kCtrlStall], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
}

// RZ_PROFILE stage stamps of the fused forms: one row of wall-clock values per
// workgroup (mark workgroups of the one-launch form first, then the slices)
enum StampIndex : int {
  kStampStart = 0,     // the workgroup started
  kStampReady = 1,     // slice: the marks of this launch are in (one-launch form)
  kStampLookback = 2,  // slice: the look-back (in the last step) has the earlier slices' orphan counts
  kStampMoves = 3,     // slice: every step's moves issued
  kStampExchange = 4,  // slice: every orphan of the slice went through its exchange word
  kStampExit = 5,      // slice: left (after clearing its bitmap words and counting its exit)
  kStampMarked = 1,    // mark workgroup: its marks are done
  kStampRaised = 2,    // mark workgroup: (the last one) raised the ready words
  kStampsPerGroup = 8
};

// internal flag: the tail bitmap was set by the caller (rz_remove_premarked)
constexpr uint32_t kFlagPremarked = 0x80000000u;

constexpr int kMaxSlices = 4096;  // cap on move-kernel workgroups (LDS prefix in the pairing kernel)

struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
};

}  // namespace rz

struct rz_ctx {
  int device = 0;
  int cu_count = 0;
  char arch[64] = {0};
  int max_wg_per_cu_256 = 0;  // occupancy of the cooperative kernel at 256 threads
  int lds_per_block = 65536;  // LDS bytes the fused form's orphan list may use (capped at 64 KiB)
  int lds_max = 65536;        // LDS bytes one workgroup may use on this device
  // scratch, all zero-initialised on allocation and kept zero by the kernels
  rz::DeviceBuffer tail_bits;  // k bits: tail removals of the plain path
  rz::DeviceBuffer list_bits;  // n bits: removals of the O(n) path and validation
  rz::DeviceBuffer counts;     // per-slice orphan counts (uint64: dest | src << 32)
  rz::DeviceBuffer orphans;    // dest orphans when RZ_PRESERVE_REMOVALS is set
  rz::DeviceBuffer status;     // look-back status words of the partition
  rz::DeviceBuffer ploaded;    // partition: per tile, the epoch once its items are loaded
  rz::DeviceBuffer pgcnt;      // partition: per 64 tiles, the number of them loaded (zero between calls)
  rz::DeviceBuffer pgdone;     // partition: per 64 tiles, the epoch once all of them are loaded
  rz::DeviceBuffer pent;       // partition's bucketed marks: per 2^19-position bin, its removals
  rz::DeviceBuffer pcur;       // partition's bucketed marks: per bin its count, then per bin a done counter (zero between calls)
  rz::DeviceBuffer ctrl;       // control block (rz::CtrlWord)
  rz::DeviceBuffer queue;      // ticket / exit / done words of the single-list queue launch
  rz::DeviceBuffer mstatus;    // fused form: two halves of kMaxSlices look-back words, used alternately
  rz::DeviceBuffer slots;      // fused form: one exchange word per orphan rank (epoch | position + 1)
  rz::DeviceBuffer dense_cnt;    // dense form: per-tile bin counts, then their per-bin prefixes
  rz::DeviceBuffer dense_small;  // dense form: bin totals / bases, per-tile orphan counts / ranks
  rz::DeviceBuffer dense_ent;    // dense form: coarse-bin head writes (positions, then values)
  rz::DeviceBuffer dense_fine;   // dense form: fine-bin head writes (fixed capacity per bin)
  rz::DeviceBuffer dense_src;    // dense form: source orphan values by rank
  rz::DeviceBuffer gstatus;    // preserve-list gather: the gathered items before they are copied back
  rz::DeviceBuffer dispatch;   // network parameters, one list descriptor, its features and the choice
  bool dispatch_ready = false; // parameters installed (else the k < n/2 rule)
  uint8_t* auto_host = nullptr;  // pinned staging of the descriptor and the choice
  rz::DeviceBuffer choose;       // rz_dispatch_choose: descriptors, features, choices
  uint8_t* choose_host = nullptr;  // its pinned staging: descriptors, then choices
  uint8_t* choose_host_dev = nullptr;  // the same staging's device address
  size_t choose_host_bytes = 0;
  hipEvent_t choose_ev = nullptr;  // recorded after the choices' copy back
  int32_t choose_n = -1;           // lists of the choice not yet collected (-1: none)
  bool choose_dev = false;         // that choice is the network's (else the host's k < n/2 rule)
  rz_tile_config batch_cfg;    // default launch shape of batches
  rz_batch* last_batch = nullptr;  // plan of the last rz_remove_batched call
  std::vector<rz_batch*> batches;  // batches made by rz_batch_create (their scratch is reset with the context's)
  uint32_t* ctrl_host = nullptr;  // pinned mirror for stats / validation reads
  rz_tile_config table[3][33];    // [element size 4/8/16][floor(log2 k)]
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};  // RZ_PROFILE marks
  bool ev_recorded = false;
  rz::DeviceBuffer stamps;     // RZ_PROFILE: stage stamps of the last fused call (StampIndex rows)
  uint32_t stamp_groups = 0;   // rows written by that call (0: the last call had none)
  uint32_t stamp_marks = 0;    // of which mark workgroups (one-launch form)
  std::mutex mu;
};

namespace rz {

rz_status ensure(rz_ctx* ctx, DeviceBuffer& b, size_t bytes, hipStream_t s);
// Makes `dev` the current device for a scope and restores the previous one.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
    else prev = -1;
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};
inline int esz_index(uint32_t esz) { return esz == 4 ? 0 : esz == 8 ? 1 : esz == 16 ? 2 : -1; }
inline int log2_floor(int64_t v) {
  int r = 0;
  while ((v >> 1) > 0) { v >>= 1; ++r; }
  return r;
}
rz_tile_config default_config(uint32_t esz, int64_t n, int64_t k);

// launchers (one per translation unit)
// ext_bits: a caller-owned tail bitmap (with kFlagPremarked: already marked)
rz_status launch_plain(rz_ctx* ctx, void* data, int64_t n, uint32_t esz, int32_t* R, int64_t k, uint32_t flags,
                       const rz_tile_config& cfg, hipStream_t s, uint32_t* ext_bits = nullptr);
rz_status launch_mark_removals(rz_ctx* ctx, const int32_t* R, int64_t n, int64_t k, uint32_t* bits, hipStream_t s);
// rz_remove_sized: the one-launch fused shape for up to k_cap removals (false
// when none fits), and its launch with n and k read from d_n, d_k on the device
bool sized_shape(const rz_ctx* ctx, uint32_t esz, int64_t k_cap, rz_tile_config* out);
rz_status launch_sized(rz_ctx* ctx, void* data, int64_t n_cap, uint32_t esz, const int32_t* R, int64_t k_cap,
                       const int32_t* d_n, const int32_t* d_k, uint32_t flags, const rz_tile_config& cfg,
                       hipStream_t s);
rz_status launch_partition(rz_ctx* ctx, void* data, int64_t n, uint32_t esz, const int32_t* R, int64_t k,
                           uint32_t flags, hipStream_t s);
rz_status launch_validate(rz_ctx* ctx, const int32_t* R, int64_t n, int64_t k, uint32_t flags, hipStream_t s);
rz_status launch_queue_single(rz_ctx* ctx, void* data, int64_t n, uint32_t esz, int32_t* R, int64_t k,
                              uint32_t flags, const rz_tile_config& cfg, hipStream_t s);
// rz_remove_batched: plan into the context's reusable batch (buffers only grow,
// the plan is uploaded asynchronously) and launch it.
rz_status batch_plan_and_launch(rz_ctx* ctx, const rz_list_desc* lists, int32_t nlists, uint32_t flags,
                                hipStream_t s);
// slices of the move step: L pairs each (a multiple of wg_size*items), ns = ceil(k/L) <= max_slices_cap
void slice_shape(int64_t k, const rz_tile_config& cfg, int max_slices_cap, uint32_t* L, uint32_t* ns);
// dense form (rz_dense.hip): every step after the tail marks
template <typename T>
rz_status dense_moves(rz_ctx* ctx, T* A, int64_t n, const int32_t* R, int64_t k, const rz_tile_config& cfg,
                      uint32_t* bits, uint32_t* ctrl, int stats, bool bucket_marks, hipStream_t s);
// dense form: whether it sets the tail marks itself (bucketed) instead of the atomic mark kernel
bool dense_bucket_marks(const rz_tile_config& cfg, int64_t k, uint32_t flags);
// Bucketed marks (rz_partition.hip): removals r in [base, base + m) set bit
// r - base of the m-bit bitmap `bits` (zero on entry); epoch_ctrl (may be
// null) gets the compaction's epoch started, stat (may be null) the number of
// removals in the window.  m up to 2^29 (bucket_marks_fit).
bool bucket_marks_fit(int64_t m);
rz_status bucket_marks(rz_ctx* ctx, const int32_t* R, int64_t k, uint32_t base, int64_t m, uint32_t* bits,
                       uint32_t* epoch_ctrl, uint64_t* stat, hipStream_t s);
// dense form: whether (cfg, n, k) can be launched (checked before any kernel of a call runs)
bool dense_shape_ok(const rz_ctx* ctx, const rz_tile_config& cfg, uint32_t esz, int64_t n, int64_t k, uint32_t flags);
// After a failed device check: zero the scratch the kernels keep zero between
// calls (context buffers, state words of the control block, batches), so that
// the next call starts clean.  Synchronises the device.
rz_status reset_scratch(rz_ctx* ctx);
rz_status batch_reset_scratch(rz_batch* b);
void batch_detach(rz_batch* b);  // the context is being destroyed
// Reads the error word into ctrl_host (all of the control block when `all`),
// waits for the stream, and returns RZ_EDEVICE (after reset_scratch) if a bit is set.
rz_status read_errors(rz_ctx* ctx, hipStream_t s, bool all, uint32_t* error_flags);
// Scratch of one fused call that the caller provides (rz_remove_noalloc): the
// control block (kCtrlWords words), two halves of kMaxSlices look-back words,
// ceil(k/32) tail bitmap words and k/2 + 1 exchange words; err_out (may be
// null) receives the call's error bits.
struct FusedScratch {
  uint32_t* ctrl;
  uint64_t* mstatus;
  uint32_t* bits;
  uint64_t* slots;
  uint32_t* err_out;
};
// The fused form (variant 4: mark kernel, then rz_move_pair; variant 6: one
// launch) on that scratch, after a prologue kernel that clears the bitmap words
// of this k, the error word and the exit counters.  Allocates nothing.
rz_status launch_fused_static(void* data, int64_t n, uint32_t esz, const int32_t* R, int64_t k, uint32_t flags,
                              const rz_tile_config& cfg, const FusedScratch& fs, hipStream_t s);
// RZ_STATS / RZ_VALIDATE: the checking read at the end of a call
rz_status finish_call(rz_ctx* ctx, uint32_t flags, hipStream_t s);
// RZ_PROFILE: record mark i (0..3) on the stream when profiling
rz_status prof_mark(rz_ctx* ctx, uint32_t flags, int i, hipStream_t s);
// fill `count` uint32 words with `v` by a kernel (graph-capture safe, unlike hipMemsetAsync)
rz_status fill_u32(uint32_t* p, uint32_t v, size_t count, hipStream_t s);

}  // namespace rz

#define RZ_HIP(call)                                  \
  do {                                                \
    hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return RZ_EHIP;             \
  } while (0)
