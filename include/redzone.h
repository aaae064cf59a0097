/*
 * redzone.h -- C ABI of the MI355X (gfx950) Redzone O(k) in-place stream
 * compaction library (librz_amd.so).
 *
 * Every pointer named d_* is device memory on the context's device.  Streams
 * are hipStream_t passed as void*.  No entry point synchronises the stream
 * unless its comment says so.  All sizes are element counts.
 *
 * Contract shared by the Redzone entry points (the reference shape
 * `(T* List, int Length, int* Removals, int RemoveCount)`, redzone.hxx:627,
 * redzone.hxx:238, Redzone_CPU.hxx:150):
 *   - d_data holds n items of elem_bytes each (4, 8 or 16); d_removals holds k
 *     distinct indices in [0, n), in any order; 0 <= k <= n < 2^31.
 *   - On return (stream order) d_data[0, n-k) holds exactly the survivors.
 *     Every survivor below n-k that was not removed keeps its index; the holes
 *     are filled from the tail [n-k, n) in the canonical order (the i-th
 *     removal takes tail item i when both are valid; the leftover holes take
 *     the leftover tail items in ascending order).  The result does not depend
 *     on the launch configuration.
 *   - d_data[n-k, n) is scratch afterwards.  d_removals is scratch afterwards
 *     unless RZ_PRESERVE_REMOVALS is set (the reference clobbers it:
 *     redzone.hxx:803, Redzone_CPU.hxx:19).
 *   - Item payloads are moved bit for bit; no payload value is reserved (the
 *     reference's GPU path reserves 0x80000000, redzone.hxx:704; this one
 *     marks a tail bitmap instead).
 *   - k == 0 and k == n return at once (redzone.hxx:651).
 */
#ifndef REDZONE_AMD_H
#define REDZONE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rz_status {
  RZ_OK = 0,
  RZ_EINVAL = 1,   /* bad argument: k > n, n >= 2^31, null pointer, element size */
  RZ_ERANGE = 2,   /* a removal index outside [0, n) (RZ_VALIDATE only) */
  RZ_EDUP = 3,     /* a duplicated removal index (RZ_VALIDATE only) */
  RZ_ECOOP = 4,    /* cooperative launch refused (grid not co-resident) */
  RZ_EHIP = 5,     /* a HIP runtime call failed */
  RZ_ENOMEM = 6,   /* scratch allocation failed */
  RZ_ENOTUNED = 7, /* no learned model loaded where one is required */
  RZ_EDEVICE = 8,  /* a check made on the device failed (see rz_check_errors): the output of the calls since the
                      last check is unspecified; the context's scratch has been reset */
  RZ_ETOOBIG = 9   /* k above the fixed scratch of rz_remove_noalloc (RZ_NOALLOC_MAX_K); nothing was run */
} rz_status;

/* flags */
#define RZ_PRESERVE_REMOVALS 0x1u /* leave d_removals unchanged (dest orphans go to context scratch) */
#define RZ_VALIDATE 0x2u          /* check range and duplicates first, and the device error word after the
                                     call (as rz_check_errors); synchronises the stream */
#define RZ_COOPERATIVE 0x4u       /* single cooperative launch with in-kernel grid barriers */
#define RZ_STATS 0x8u             /* record per-call statistics readable with rz_last_stats, and check the
                                     device error word after the call (as rz_check_errors); synchronises */
#define RZ_PROFILE 0x10u          /* time each kernel with events on the stream; read with rz_last_profile */
#define RZ_SORTED 0x20u           /* d_removals is strictly ascending: the mark step reads only its red-zone suffix
                                     (checked by RZ_VALIDATE; if it is not true the result is unspecified) */
#define RZ_ANY_PAIRING 0x40u      /* the fused forms (variants 4 and 6) may pair orphans inside a slice first
                                     (paper p.14, early pairing): the same survivors, every kept item outside the
                                     tail still in place, but which tail item fills which hole may differ from
                                     the canonical array (redzone.hxx:913-934 pairs by global rank).  Other forms
                                     return the canonical array, which also meets this contract.  With RZ_STATS
                                     the orphan count is that of the pairs left to the global exchange. */

typedef struct rz_ctx rz_ctx;

/* Launch shape.  A step is wg_size * items pairs; the removal list is cut into
 * slices of whole steps, one workgroup (or queue item) per slice. */
typedef struct rz_tile_config {
  int32_t wg_size;     /* 256 or 512 */
  int32_t items;       /* items per thread per step: 2, 4, 8, 16 (variants 0/1); 4 or 8 (variant 2, batches) */
  int32_t max_slices;  /* variants 0/1/3/4/6: upper bound on slices (workgroups of the move step);
                          batches: queues in use, 1, 2 or 4 (any other value: one per XCD, 8); 16: 8 queues
                          with every list's items dealt over all of them */
  int32_t variant;     /* 0 = three kernels on the stream (mark, move, pair), 1 = one cooperative kernel,
                          2 = persistent queue, 3 = one ordered launch (roles by workgroup id),
                          (batches take 2 = MARK, MOVE and PAIR items per list, 4 = MARK and fused
                          items: a list of several slices of at most 8,192 pairs each pairs its orphans
                          inside the move items, through look-back and exchange words, or 5 = as 2, with
                          the tail marks of lists of 2^19 to 2^22 removals grouped in buckets of 2^15
                          tail positions and each bucket's bitmap words written whole),
                          4 = two kernels (mark, then move with in-kernel orphan pairing; slices longer than
                              the workgroup LDS allows run as variant 0),
                          5 = dense form: head writes grouped by destination bin, then one scatter per bin
                              (for k/n of several percent and up),
                          6 = variant 4 in one launch: mark workgroups first, move slices after them wait for
                              the marks (same LDS limit and fallback as variant 4) */
  int32_t slice_steps; /* variant 2 and batches: steps per slice (0 = 4); variant 5: log2 of the target
                          number of destination bins, 5..10 (0 = 9); variant 6: log2 of the removals
                          each mark workgroup takes, 8..16 (0 = 12) */
  int32_t chunk;       /* variant 2 and batches: queue items taken per ticket (0 = 1) */
} rz_tile_config;

/* Removal with the tail marks already set by the caller's producer of the
 * list (see redzone_device.h: rz_mark_removal).  d_tail_bits holds
 * ceil(k/32) words with bit r-(n-k) set for every removal r >= n-k and no
 * other bit; the call skips the mark step and leaves the bitmap zero.  cfg may
 * be null (the tuned table); shapes whose single launch includes the mark
 * (variants 1-3 and 6) run as variant 4.  RZ_SORTED is ignored and RZ_STATS counts
 * no tail removals (the marks were not made here).  Has no counterpart in the
 * reference: its phase 1 always scans the list (redzone.hxx:698-707). */
rz_status rz_remove_premarked(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals,
                              int64_t k, uint32_t* d_tail_bits, uint32_t flags, const rz_tile_config* cfg,
                              void* stream);
/* Add and remove in one call (paper p.9 sec. 4.6: append G, then Redzone on
 * A+G): the g items d_add are appended after the n items of d_data (which must
 * hold capacity >= n + g items), then the k removals (indices into the
 * combined list, usually < n) are removed as by rz_remove on n + g items.  The
 * new length is n + g - k.  When g >= k and every removal is below n, the red
 * zone lies inside the appended items, so every removal is a plain move of an
 * added item into its hole. */
rz_status rz_add_remove(rz_ctx* ctx, void* d_data, int64_t n, int64_t capacity, uint32_t elem_bytes,
                        const void* d_add, int64_t g, int32_t* d_removals, int64_t k, uint32_t flags, void* stream);
/* In-place gather by a preserve list (paper p.14, eq. 3): A'[i] = A[P[i]] for
 * i < p, with P strictly ascending and P[p-1] < n (checked with RZ_VALIDATE).
 * The survivors keep their order.  From p = 2^22 on, one in-place pass (tiles
 * write in launch order once all earlier tiles have read their sources; no
 * scratch for the items); below, a gather into context scratch (p items) and
 * a copy back. */
rz_status rz_gather_preserve(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, const int32_t* d_keep,
                             int64_t p, uint32_t flags, void* stream);
/* The mark step alone, into a caller-owned tail bitmap (zero beforehand),
 * e.g. on another stream while the list is still being written; built on
 * rz_mark_removal.  redzone.hxx:698-707 without the data sentinel. */
rz_status rz_mark_removals(rz_ctx* ctx, const int32_t* d_removals, int64_t n, int64_t k, uint32_t* d_tail_bits,
                           void* stream);

/* Statistics of the last call made with RZ_STATS. */
typedef struct rz_stats {
  int64_t n, k;
  int64_t tail_removals;  /* k_z = |{r in R : r >= n-k}| */
  int64_t orphan_pairs;   /* dest orphans (== source orphans) */
  int64_t slices;
  int32_t error_flags;    /* the device error word read by that call (RZ_ERR_* bits; 0 = none) */
  int32_t pad;
} rz_stats;

/* Bits of the device error word.  The kernels OR them in; they stay set until
 * a checking call (rz_check_errors, or a call with RZ_STATS or RZ_VALIDATE)
 * reads and clears the word. */
#define RZ_ERR_ORPHAN_MISMATCH 0x1u /* dest and source orphan totals differ: the removal list broke the contract
                                       (a duplicate or an index out of range); the reference asserts the same
                                       equality in debug builds, redzone.hxx:905 */
#define RZ_ERR_WAIT_TIMEOUT 0x2u    /* an in-kernel wait on other workgroups (grid barrier, mark count, ready
                                       word) gave up: workgroups were not dispatched as the launch assumes */
#define RZ_ERR_LOOKBACK_TIMEOUT 0x4u /* a decoupled look-back gave up, for the same reason */
#define RZ_ERR_SIZES 0x8u           /* rz_remove_sized: the device-resident n and k broke 0 <= k <= n <= n_cap,
                                       k <= k_cap; the call moved nothing */

/* ---- context ------------------------------------------------------------ */
/* The context owns all device scratch (tail bitmap, slice counts, look-back
 * state, barrier words).  One context per stream at a time; calls that share a
 * context must be ordered (same stream, or synchronised).  No counterpart in
 * the reference, whose scratch BlockStorage_t is allocated by a host
 * constructor and never freed (redzone.hxx:26-81). */
rz_status rz_ctx_create(rz_ctx** out, int32_t device);
rz_status rz_ctx_destroy(rz_ctx* ctx);
const char* rz_status_string(rz_status s);
/* Waits for `stream`, then reads and clears the device error word of the
 * context (the checks the kernels make: orphan totals, and every bounded wait
 * on other workgroups; the reference's counterpart is the debug assert of
 * redzone.hxx:905, which stops the program).  *error_flags (may be null) gets
 * the RZ_ERR_* bits.  Returns RZ_OK when no bit was set.  Otherwise every call
 * on the context since the previous check may have produced a wrong array: the
 * scratch the kernels keep zero between calls is zeroed again (this
 * synchronises the device) and RZ_EDEVICE is returned.  Errors of calls on
 * other streams are seen once those calls have finished. */
rz_status rz_check_errors(rz_ctx* ctx, void* stream, uint32_t* error_flags);
/* Test hook for the failure path: the bounded in-kernel waits give up after
 * `spin_cap` polls (0 = their own limits), and with stall != 0 the waits of
 * the Redzone forms (look-back, ready word, counters) never see their
 * condition met, so each of them gives up and sets its RZ_ERR_* bit; every
 * kernel still runs to its end and stays inside its buffers.  Synchronous.
 * No counterpart in the reference. */
rz_status rz_debug_set_waits(rz_ctx* ctx, uint32_t spin_cap, uint32_t stall);
/* Device properties seen by the context: compute units and HBM peak (GB/s) used for reports. */
rz_status rz_device_info(rz_ctx* ctx, int32_t* cu_count, char* arch, int32_t arch_len);

/* ---- plain Redzone (one list) ------------------------------------------- */
/* Replaces RemoveFromList_GPUWide (redzone.hxx:626-941, launched from
 * test_redzone at redzone.hxx:1145) and RemoveFromList2 (redzone.hxx:236-444):
 * O(k) in-place removal.  Launch shape from the context's tuning table. */
rz_status rz_remove(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals, int64_t k,
                    uint32_t flags, void* stream);
/* Typed forms of rz_remove for 4-, 8- and 16-byte items. */
rz_status rz_remove_u32(rz_ctx* ctx, uint32_t* d_data, int64_t n, int32_t* d_removals, int64_t k, uint32_t flags,
                        void* stream);
rz_status rz_remove_u64(rz_ctx* ctx, uint64_t* d_data, int64_t n, int32_t* d_removals, int64_t k, uint32_t flags,
                        void* stream);
rz_status rz_remove_rec16(rz_ctx* ctx, void* d_data, int64_t n, int32_t* d_removals, int64_t k, uint32_t flags,
                          void* stream);
/* rz_remove with an explicit launch shape (autotuner and tests). */
rz_status rz_remove_config(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals,
                           int64_t k, uint32_t flags, const rz_tile_config* cfg, void* stream);
/* rz_remove with n and k in device memory, as the reference's launch wrapper
 * takes them (TestGPUWide(int* A, int* sizeA, int* R, int* sizeR, ...) reads
 * *sizeA and *sizeR inside the kernel, redzone.hxx:457-465): d_n and d_k each
 * point at one int32, written before the call in stream order (by a kernel
 * that decides k, for instance), so no host round trip is needed, and a
 * captured graph of the call removes whatever k the device holds at each
 * replay.  n_cap bounds n (d_data holds n_cap items) and k_cap bounds k (d_removals
 * holds k_cap entries, the first k of which are read); the launch and the
 * context scratch are sized for k_cap, and the workgroups beyond what the
 * device k needs leave at once.  The result is rz_remove's for (n, k).  Device
 * values outside 0 <= k <= n <= n_cap, k <= k_cap move nothing and set
 * RZ_ERR_SIZES (read by rz_check_errors).  Runs the one-launch fused form
 * (variant 6) with the tuned shape for k_cap when that is a fused form, else
 * 256 x 8; RZ_EINVAL if no fused shape fits k_cap (more than about 58.7 M
 * removals: use rz_remove).  Flags: RZ_PRESERVE_REMOVALS, RZ_SORTED,
 * RZ_ANY_PAIRING, RZ_STATS, RZ_PROFILE; RZ_VALIDATE (which needs k on the
 * host) and RZ_COOPERATIVE are RZ_EINVAL.  A graph capture needs an eager call
 * with the same caps first, which grows the context scratch. */
rz_status rz_remove_sized(rz_ctx* ctx, void* d_data, int64_t n_cap, uint32_t elem_bytes, int32_t* d_removals,
                          int64_t k_cap, const int32_t* d_n, const int32_t* d_k, uint32_t flags, void* stream);
/* Reads the statistics recorded by the last call made with RZ_STATS. */
rz_status rz_last_stats(rz_ctx* ctx, rz_stats* out);
/* Kernel times (ms) of the last call made with RZ_PROFILE: ms[0] mark, ms[1]
 * move (or the whole cooperative kernel / the partition pass), ms[2] orphan
 * pairing, ms[3] first kernel start to last kernel end.  Waits for the call. */
rz_status rz_last_profile(rz_ctx* ctx, float ms[4]);
/* Stage timeline (us) of the last fused-form call (variants 4 and 6) made with
 * RZ_PROFILE: every workgroup stamps the wall clock at its stage boundaries;
 * times are from the first workgroup's start.  The counterpart of the
 * reference's per-phase clock64() stamps (redzone.hxx:629, 708-712, 847-851,
 * 936-939).  out[0] last mark workgroup done, out[1] ready words raised,
 * out[2] / out[3] first / last slice started, then the last slice to finish
 * each stage: out[4] ready wait, out[5] look-back (made by the last step
 * before that step's moves), out[6] moves, out[7] orphan exchange, out[8]
 * exit; out[9..13] the mean time a slice spends in each of those five stages
 * (from the end of the one before); out[14] mark workgroups, out[15] slices.
 * Synchronises the device.  RZ_EINVAL if the last call had no stamps. */
rz_status rz_last_stages(rz_ctx* ctx, float out[16]);

/* ---- O(n) stable partition ---------------------------------------------- */
/* Replaces the O(n) comparators: remove_phase1 (DataStruct.hxx:11-18) followed
 * by StreamCompact_OneBlock (redzone.hxx:526-603) / thrust::remove
 * (Thrust_remove.hxx:15-17) / std::remove (std_parallel_remove.hxx:34-36).
 * Marks the removals in a bitmap, then compacts all n items in place in one
 * decoupled look-back pass.  Survivors keep their relative order.  d_removals
 * is not modified. */
rz_status rz_partition_on(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, const int32_t* d_removals,
                          int64_t k, uint32_t flags, void* stream);

/* ---- batched Redzone ----------------------------------------------------- */
/* No counterpart in the reference (it removes from one list per launch).
 * Many independent lists are compacted by one persistent kernel that pulls
 * (list, slice) work items from device queues, one per XCD (each list's items
 * in one queue, list after list); every list gets exactly the result
 * rz_remove would give it.  Lists must not overlap. */
typedef struct rz_list_desc {
  void* data;          /* device: n items of elem_bytes each */
  int32_t* removals;   /* device: k distinct indices in [0, n) */
  int64_t n;
  int64_t k;
  uint32_t elem_bytes; /* 4, 8 or 16 */
  uint32_t method;     /* RZ_METHOD_PARTITION (3): the list runs the O(n) stable partition inside the batch
                          (survivors in their original order, as rz_partition_on); any other value: Redzone */
} rz_list_desc;

typedef struct rz_batch rz_batch;

typedef struct rz_batch_info {
  int64_t nitems;        /* work items in the queues */
  int64_t bitmap_bytes;  /* tail bitmap scratch */
  int32_t grid;          /* persistent workgroups launched */
  int32_t lists_active;  /* lists with 0 < k < n */
} rz_batch_info;

/* Plans a batch from a HOST array of descriptors (pointers are device
 * pointers) and allocates its scratch; uploads the plan synchronously.  The
 * batch can then be launched any number of times on the same buffers (new
 * data, same shapes).  cfg may be null (context default). */
rz_status rz_batch_create(rz_ctx* ctx, const rz_list_desc* lists, int32_t nlists, uint32_t flags,
                          const rz_tile_config* cfg, rz_batch** out);
/* One persistent launch (or a replay of the captured graph); no allocation,
 * no memset, so it may itself be captured by the caller.  RZ_EINVAL once the
 * batch's context has been destroyed (the batch can still be destroyed). */
rz_status rz_batch_launch(rz_batch* b, void* stream);
/* Captures the launch into a hipGraph owned by the batch; later
 * rz_batch_launch calls replay it. */
rz_status rz_batch_capture(rz_batch* b, void* stream);
rz_status rz_batch_get_info(rz_batch* b, rz_batch_info* out);
/* Waits for the device, then frees the batch. */
rz_status rz_batch_destroy(rz_batch* b);
/* One-shot: plan + launch on a batch owned by the context.  Its device
 * buffers only grow and the plan is uploaded asynchronously on the stream, so
 * repeated calls neither allocate nor synchronise in steady state. */
rz_status rz_remove_batched(rz_ctx* ctx, const rz_list_desc* lists, int32_t nlists, uint32_t flags, void* stream);

/* ---- synthetic inputs ---------------------------------------------------- */
/* Device generators of benchmark and tuning inputs.  No counterpart in the
 * reference, whose generator (TestLists_t::Update, DataStruct.hxx:73-133)
 * draws on the host.  rz_gen_uniform_removals writes k distinct indices in
 * [0, n): the first k images of a keyed bijection of [0, n), in that order. */
rz_status rz_gen_uniform_removals(rz_ctx* ctx, int32_t* d_removals, int64_t n, int64_t k, uint64_t seed,
                                  void* stream);
/* n items of elem_bytes: the first 4-byte word of item i is a keyed bijection
 * of i (unique), the other words are hashes of i. */
rz_status rz_gen_payload(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, uint64_t seed, void* stream);

/* ---- autotuner ----------------------------------------------------------- */
/* No counterpart in the reference (its launch shape is fixed at compile time:
 * GPUBlockCount / GPUBlockSize, tests.cu:1-2).  The context keeps one launch
 * shape per (element size, floor(log2 k)) for rz_remove and one for batches. */
/* Built-in candidate lists (batch = 0: plain path, 1: batches). */
rz_status rz_tune_candidates(int32_t batch, rz_tile_config* out, int32_t max, int32_t* count);
/* Times every candidate (null = built-in list) on a synthetic list of n items
 * with k uniform removals (median of reps calls; cold = read a 1 GiB scrub
 * buffer before each call) and installs the fastest for (elem_bytes, log2 k).
 * A candidate whose survivors differ from the three-kernel form's, or that
 * sets a device error bit, is rejected: all_ms gets -2 for it (-1 for a shape
 * the device refuses).  Allocates and frees its own inputs; synchronises. */
rz_status rz_tune(rz_ctx* ctx, uint32_t elem_bytes, int64_t n, int64_t k, const rz_tile_config* candidates,
                  int32_t ncand, int32_t reps, int32_t cold, rz_tile_config* best, float* best_ms, float* all_ms);
/* The same for batches: nlists lists, n log-uniform in [n_min, n_max],
 * k = floor(r n) with r uniform in [0, r_max]; every list's survivors are
 * checked against the three-kernel form's. */
rz_status rz_tune_batch(rz_ctx* ctx, uint32_t elem_bytes, int32_t nlists, int64_t n_min, int64_t n_max,
                        double r_max, int32_t reps, const rz_tile_config* candidates, int32_t ncand,
                        rz_tile_config* best, float* best_ms, float* all_ms);
rz_status rz_table_get(rz_ctx* ctx, uint32_t elem_bytes, int32_t log2k, rz_tile_config* out);
rz_status rz_table_set(rz_ctx* ctx, uint32_t elem_bytes, int32_t log2k, const rz_tile_config* cfg);
rz_status rz_batch_config_get(rz_ctx* ctx, rz_tile_config* out);
rz_status rz_batch_config_set(rz_ctx* ctx, const rz_tile_config* cfg);
/* Saves / loads the whole table as text ("plain esz log2k wg items max_slices
 * variant slice_steps chunk" and "batch wg items max_slices variant
 * slice_steps chunk" lines). */
rz_status rz_table_save(rz_ctx* ctx, const char* path);
rz_status rz_table_load(rz_ctx* ctx, const char* path);

/* ---- learned dispatch ---------------------------------------------------- */
/* No counterpart in the reference, which states the fixed rule "Redzone is
 * faster than O(n) when k < n/2" (README.md:6; paper p.14).  A small FP32 MLP
 * predicts the log2 time of each method from features of a list; the
 * dispatcher runs the method with the smallest prediction.  Layers F -> H ->
 * H -> H -> C with ReLU; parameters are row-major W[out][in] then b[out] per
 * layer, in layer order (RZ_MLP_PARAMS floats). */
#define RZ_MLP_FEATURES 12
#define RZ_MLP_HIDDEN 32
#define RZ_MLP_OUTPUTS 4
#define RZ_MLP_PARAMS 2660
/* Features of each list (device array of descriptors) into d_features
 * [nlists][RZ_MLP_FEATURES]: log2 n / 32, log2 k / 32, k/n, log2(k/n) / 20,
 * elem_bytes / 16, and from a 256-entry strided sample of the removal list:
 * fraction in the red zone, mean and std of position / n, fraction of adjacent
 * entries that are consecutive / within 64, fraction ascending; then 1. */
rz_status rz_features(rz_ctx* ctx, const rz_list_desc* d_lists, int32_t nlists, float* d_features, void* stream);
/* Forward pass: d_y [batch][RZ_MLP_OUTPUTS] (may be null) and the argmin per
 * row in d_choice (may be null). */
rz_status rz_mlp_forward(rz_ctx* ctx, const float* d_params, const float* d_x, int32_t batch, float* d_y,
                         int32_t* d_choice, void* stream);
/* Gradient of the masked mean squared error sum m (y - t)^2 / mask_sum over
 * the batch into d_grad [RZ_MLP_PARAMS] and the loss into d_loss (may be
 * null).  d_work holds batch * (RZ_MLP_PARAMS + 1) floats.  Deterministic. */
rz_status rz_mlp_grad(rz_ctx* ctx, const float* d_params, const float* d_x, const float* d_target,
                      const float* d_mask, int32_t batch, float mask_sum, float* d_grad, float* d_loss,
                      float* d_work, void* stream);
/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with step >= 1. */
rz_status rz_mlp_adam(rz_ctx* ctx, float* d_params, const float* d_grad, float* d_m, float* d_v, int32_t step,
                      float lr, void* stream);

/* Methods of the dispatcher (outputs of the network, in order). */
#define RZ_METHOD_REDZONE 0   /* rz_remove with the tuned launch shape */
#define RZ_METHOD_ORDERED 1   /* one ordered launch (variant 3) */
#define RZ_METHOD_DENSE 2     /* the dense form (variant 5) */
#define RZ_METHOD_PARTITION 3 /* rz_partition_on */
/* Launch shape of RZ_METHOD_ORDERED / RZ_METHOD_DENSE for an element size
 * (RZ_EINVAL for the other methods, which have none).  Needs no context. */
rz_status rz_dispatch_method_config(int32_t method, uint32_t elem_bytes, rz_tile_config* out);
/* Installs the network's parameters (host array of RZ_MLP_PARAMS floats) in
 * the context; null removes them, and rz_remove_auto then uses the k < n/2
 * rule.  Synchronous. */
rz_status rz_dispatch_set_params(rz_ctx* ctx, const float* h_params);
/* The reference's remove(data, n, removeList, k) with the method chosen per
 * list: the features of the list and the network's argmin are computed on
 * the stream, the choice is read back (this synchronises the stream), then
 * the chosen method runs on the stream.  *method (may be null) gets the
 * RZ_METHOD_* used.  Replaces the fixed choice the reference states
 * (README.md:6; paper p.14: Redzone below k = n/2, O(n) above). */
rz_status rz_remove_auto(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals, int64_t k,
                         uint32_t flags, int32_t* method, void* stream);
/* The network's method for many lists, issued without waiting (no counterpart
 * in the reference, whose fixed rule is README.md:6).  rz_dispatch_choose
 * copies the host descriptors h_lists (method fields ignored) to the device
 * through pinned staging, computes their features and the forward pass on the
 * stream (in stream order after the removal lists' producers), copies the
 * choices back into pinned staging and records an event, then returns.
 * d_params: RZ_MLP_PARAMS device floats, or null for the parameters installed
 * by rz_dispatch_set_params; with neither, the k < n/2 rule applies.
 * rz_dispatch_collect waits for that event only and writes one RZ_METHOD_* per
 * list into h_choice (nlists must match the last rz_dispatch_choose); lists
 * with k = 0 or k = n get RZ_METHOD_REDZONE, a value out of range the rule.
 * One choice per context is outstanding at a time: a new rz_dispatch_choose
 * waits for the previous one's copy and replaces it.  Issued ahead of other
 * host work (a batch's plan), the choice is ready when that work is done. */
rz_status rz_dispatch_choose(rz_ctx* ctx, const float* d_params, const rz_list_desc* h_lists, int32_t nlists,
                             void* stream);
rz_status rz_dispatch_collect(rz_ctx* ctx, int32_t* h_choice, int32_t nlists);
/* Many lists of any sizes with the method chosen per list (no counterpart in
 * the reference; it generalises rz_remove_auto to a stream of lists): lists
 * with fewer than batch_max_k removals run in one persistent batch
 * (rz_remove_batched, their method fields ignored), the others each get their
 * own launch with the network's method.  In stream order: the choice for the
 * large lists is issued first (rz_dispatch_choose with d_params, or the
 * installed network when null, or the k < n/2 rule without either), then the
 * batch is planned and launched while the device computes it, then the
 * choice is collected (waiting for that choice only) and the large lists are
 * queued behind the batch.  methods (may be null) gets the RZ_METHOD_* that
 * ran per list (RZ_METHOD_REDZONE for the batched ones).  flags apply to
 * every call; the first failing call's status is returned.  The mixed-stream
 * measurement runs this with batch_max_k = 2^22. */
rz_status rz_remove_auto_batched(rz_ctx* ctx, const float* d_params, const rz_list_desc* lists, int32_t nlists,
                                 int64_t batch_max_k, uint32_t flags, int32_t* methods, void* stream);

/* ---- allocation-free removal -------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): rz_remove's contract on the current device with no
 * context and no allocation, freeing or host synchronisation.  All scratch (a
 * control block, look-back words, the k-bit tail bitmap and k/2 + 1 exchange
 * words) is one fixed device array of about 16.6 MiB that every GPU the
 * library runs on holds until the process ends, so k is limited to
 * RZ_NOALLOC_MAX_K (RZ_ETOOBIG above it, before anything runs).  The launch is
 * the fused form with a fixed shape by k (a prologue kernel, then the
 * one-launch form below 2^19 removals, else the mark kernel and the fused
 * move-and-pair kernel).  flags: RZ_PRESERVE_REMOVALS (always true here: the
 * fused form never writes d_removals), RZ_SORTED, RZ_ANY_PAIRING; any other
 * flag is RZ_EINVAL.  d_error_flags (may be null) is a caller-owned device
 * word into which the call ORs its RZ_ERR_* bits, in stream order.  Calls
 * that overlap on one GPU (other streams or threads) race on the shared
 * array: order them. */
#define RZ_NOALLOC_MAX_K (1 << 22)
rz_status rz_remove_noalloc(void* d_data, int64_t n, uint32_t elem_bytes, int32_t* d_removals, int64_t k,
                            uint32_t flags, uint32_t* d_error_flags, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* REDZONE_AMD_H */
