/* redzone_device.h -- device-side helper for kernels that produce removal lists.
 *
 * The mark step of Redzone (redzone.hxx:698-707) only needs, for every removal
 * r inside the red zone [n-k, n), one bit r-(n-k) of a k-bit tail bitmap.  A
 * kernel that writes the removal list can set those bits as it goes; the list
 * is then removed with rz_remove_premarked, which skips the mark step.  (The
 * paper notes that phase 1 can hide inside the producer of the list, p.7 and
 * p.14.)
 *
 * The bitmap is caller-owned device memory of RZ_TAIL_BITMAP_WORDS(k) words,
 * all zero before the producer runs; rz_remove_premarked leaves it zero again.
 */
#ifndef REDZONE_AMD_DEVICE_H
#define REDZONE_AMD_DEVICE_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define RZ_TAIL_BITMAP_WORDS(k) (((k) + 31) / 32)

/* Marks removal r of a list of n items with k removals.  Any thread of any
 * kernel ordered before rz_remove_premarked on the same stream may call it;
 * each removal once (duplicates are harmless to the bitmap but invalid input
 * to the removal). */
__device__ inline void rz_mark_removal(uint32_t* bits, int32_t r, int64_t n, int64_t k) {
  const int64_t z = (int64_t)r - (n - k);
  if (z >= 0 && z < k)
    __hip_atomic_fetch_or(&bits[z >> 5], 1u << (z & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

#endif /* REDZONE_AMD_DEVICE_H */
