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
  RZ_ENOTUNED = 7  /* no learned model loaded where one is required */
} rz_status;

/* flags */
#define RZ_PRESERVE_REMOVALS 0x1u /* leave d_removals unchanged (dest orphans go to context scratch) */
#define RZ_VALIDATE 0x2u          /* check range and duplicates first; synchronises the stream */
#define RZ_COOPERATIVE 0x4u       /* single cooperative launch with in-kernel grid barriers */
#define RZ_STATS 0x8u             /* record per-call statistics readable with rz_last_stats (synchronises) */
#define RZ_PROFILE 0x10u          /* time each kernel with events on the stream; read with rz_last_profile */

typedef struct rz_ctx rz_ctx;

/* Launch shape of the plain path: workgroup size, items per thread and the
 * cap on the number of removal slices (one slice per workgroup). */
typedef struct rz_tile_config {
  int32_t wg_size;     /* 256 or 512 */
  int32_t items;       /* items per thread per step: 2, 4, 8 or 16 */
  int32_t max_slices;  /* upper bound on workgroups of the move kernel */
  int32_t variant;     /* 0 = three kernels on the stream, 1 = one cooperative kernel */
} rz_tile_config;

/* Statistics of the last call made with RZ_STATS. */
typedef struct rz_stats {
  int64_t n, k;
  int64_t tail_removals;  /* k_z = |{r in R : r >= n-k}| */
  int64_t orphan_pairs;   /* dest orphans (== source orphans) */
  int64_t slices;
  int32_t error_flags;    /* nonzero: input violated the contract (seen on device) */
  int32_t pad;
} rz_stats;

/* ---- context ------------------------------------------------------------ */
/* The context owns all device scratch (tail bitmap, slice counts, look-back
 * state, barrier words).  One context per stream at a time; calls that share a
 * context must be ordered (same stream, or synchronised).  No counterpart in
 * the reference, whose scratch BlockStorage_t is allocated by a host
 * constructor and never freed (redzone.hxx:26-81). */
rz_status rz_ctx_create(rz_ctx** out, int32_t device);
rz_status rz_ctx_destroy(rz_ctx* ctx);
const char* rz_status_string(rz_status s);
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
/* Reads the statistics recorded by the last call made with RZ_STATS. */
rz_status rz_last_stats(rz_ctx* ctx, rz_stats* out);
/* Kernel times (ms) of the last call made with RZ_PROFILE: ms[0] mark, ms[1]
 * move (or the whole cooperative kernel / the partition pass), ms[2] orphan
 * pairing, ms[3] first kernel start to last kernel end.  Waits for the call. */
rz_status rz_last_profile(rz_ctx* ctx, float ms[4]);

/* ---- O(n) stable partition ---------------------------------------------- */
/* Replaces the O(n) comparators: remove_phase1 (DataStruct.hxx:11-18) followed
 * by StreamCompact_OneBlock (redzone.hxx:526-603) / thrust::remove
 * (Thrust_remove.hxx:15-17) / std::remove (std_parallel_remove.hxx:34-36).
 * Marks the removals in a bitmap, then compacts all n items in place in one
 * decoupled look-back pass.  Survivors keep their relative order.  d_removals
 * is not modified. */
rz_status rz_partition_on(rz_ctx* ctx, void* d_data, int64_t n, uint32_t elem_bytes, const int32_t* d_removals,
                          int64_t k, uint32_t flags, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* REDZONE_AMD_H */
