/*
 * rz_oracle.c -- CPU oracle for Redzone O(k) in-place stream compaction.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file.  Only the test suite, the smoke check and the CPU-baseline
 * leg of the benchmark use it, and only as the checker / timed CPU baseline.
 *
 * Contents
 *   rzo_canonical      sequential canonical Redzone on any element size
 *                      (paper Alg. 1, p.4, with in-order orphan pairing; the
 *                      phases follow redzone.hxx:698-707 (mark), :777-816
 *                      (move / collect orphans) and :913-934 (pair orphans)).
 *   rzo_redzone_cpu    restatement of redzone_CPU (Redzone_CPU.hxx:150-248):
 *                      T slices of stride ceil(k/T), high-bit marks in the
 *                      removal list (Redzone_CPU.hxx:12-24), in-slice orphans
 *                      (:26-59), serial exclusive prefix (:207-219), and a
 *                      load-balanced orphan pairing (:72-129).  OpenMP replaces
 *                      oneTBB; the search for the first row is a binary search
 *                      (the reference's comment at :85 asks for one).
 *   rzo_remove_stable  the O(n) comparator: mark every removal, then a stable
 *                      remove (std_parallel_remove.hxx:13-36,
 *                      DataStruct.hxx:11-18).
 *   rzo_check          invariant checker (length, survivor multiset with
 *                      payload bits, in-place invariant, optional exact order).
 *
 * Planted mistakes (mutation > 0) exist so that the test suite can show its
 * checks catch plausible errors; mutation 0 is the faithful oracle.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <omp.h>

#define RZO_OK 0
#define RZO_EINVAL (-1)
#define RZO_ERANGE (-2)
#define RZO_EDUP (-3)
#define RZO_ENOMEM (-4)

static int64_t now_ns(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (int64_t)ts.tv_sec * 1000000000LL + ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Sequential canonical Redzone.                                            */
/* ------------------------------------------------------------------------ */

/* Mutations: 1 = pair orphans in reverse order, 2 = ignore tail marks,
 * 3 = red zone starts one item late, 4 = source orphans taken from Z[i+1]. */
int rzo_canonical(void* data, int64_t n, uint32_t esz, const int32_t* R, int64_t k, int mutation) {
  if (n < 0 || k < 0 || k > n || esz == 0) return RZO_EINVAL;
  if (k == 0 || k == n) return RZO_OK; /* no-ops (redzone.hxx:651) */
  char* A = (char*)data;
  const int64_t sz = n - k + (mutation == 3 ? 1 : 0);
  const int64_t nz = n - sz; /* tail length */
  uint8_t* dead = (uint8_t*)calloc((size_t)k, 1);
  int64_t* dst_orph = (int64_t*)malloc(sizeof(int64_t) * (size_t)k);
  int64_t* src_orph = (int64_t*)malloc(sizeof(int64_t) * (size_t)k);
  if (!dead || !dst_orph || !src_orph) { free(dead); free(dst_orph); free(src_orph); return RZO_ENOMEM; }
  /* phase 1: dead[i] <=> (sz + i) is in R */
  for (int64_t i = 0; i < k; ++i) {
    int64_t r = R[i];
    if (r < 0 || r >= n) { free(dead); free(dst_orph); free(src_orph); return RZO_ERANGE; }
    if (r >= sz && mutation != 2) {
      if (dead[r - sz]) { free(dead); free(dst_orph); free(src_orph); return RZO_EDUP; }
      dead[r - sz] = 1;
    }
  }
  /* phase 2: pair i moves Z[i] into hole R[i]; unmatched halves become orphans */
  int64_t nd = 0, ns = 0;
  for (int64_t i = 0; i < k && i < nz; ++i) {
    const int64_t d = R[i];
    const int64_t zi = (mutation == 4 && i + 1 < nz) ? i + 1 : i;
    const int src_ok = !dead[i];
    const int dst_ok = d < sz;
    if (dst_ok && src_ok) memmove(A + d * esz, A + (sz + zi) * esz, esz);
    else if (!dst_ok && src_ok) src_orph[ns++] = sz + zi;
    else if (dst_ok && !src_ok) dst_orph[nd++] = d;
  }
  /* phase 3: t-th dest orphan gets the t-th source orphan */
  const int64_t m = nd < ns ? nd : ns;
  for (int64_t t = 0; t < m; ++t) {
    const int64_t s = (mutation == 1) ? src_orph[m - 1 - t] : src_orph[t];
    memmove(A + dst_orph[t] * esz, A + s * esz, esz);
  }
  free(dead); free(dst_orph); free(src_orph);
  return (nd == ns) ? RZO_OK : RZO_EDUP;
}

/* Number of removals that fall in the red zone [n-k, n). */
int64_t rzo_count_tail(const int32_t* R, int64_t n, int64_t k) {
  int64_t c = 0;
  for (int64_t i = 0; i < k; ++i) c += (R[i] >= n - k);
  return c;
}

/* ------------------------------------------------------------------------ */
/* Redzone_CPU restatement (OpenMP, T slices).                              */
/* ------------------------------------------------------------------------ */

#define HIGH 0x80000000u

static inline uint32_t ld_relaxed(const int32_t* p) { return (uint32_t)__atomic_load_n(p, __ATOMIC_RELAXED); }

/* first row r in [0, T) with prefix[r] <= idx < prefix[r+1]; rows may be empty */
static inline int find_row(const int64_t* prefix, int T, int64_t idx) {
  int lo = 0, hi = T - 1;
  while (lo < hi) { /* largest r with prefix[r] <= idx, skipping empty rows at the end */
    int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= idx) lo = mid; else hi = mid - 1;
  }
  return lo;
}

#define DEFINE_REDZONE_CPU(NAME, T_)                                                          \
static void NAME(T_* list, int64_t n, int32_t* R, int64_t k, int T, int64_t times[4]) {       \
  const int64_t sz = n - k;                                                                   \
  const int64_t stride = (k + T - 1) / T;                                                     \
  int64_t* dcnt = (int64_t*)calloc((size_t)T + 1, sizeof(int64_t));                           \
  int64_t* scnt = (int64_t*)calloc((size_t)T + 1, sizeof(int64_t));                           \
  const int64_t t0 = now_ns();                                                                \
  int64_t t1 = 0, t2 = 0;                                                                     \
  _Pragma("omp parallel num_threads(T)")                                                      \
  {                                                                                           \
    const int tid = omp_get_thread_num();                                                     \
    const int nth = omp_get_num_threads();                                                    \
    /* slices are indexed by tid; with fewer threads than asked each one takes several */    \
    for (int s = tid; s < T; s += nth) {                                                      \
      const int64_t a = stride * s, b = (a + stride < k) ? a + stride : k;                    \
      for (int64_t i = a; i < b; ++i) { /* phase 1: high-bit mark in R */                     \
        const int64_t z = (int64_t)(ld_relaxed(&R[i]) & ~HIGH) - sz;                          \
        if (z >= 0) __atomic_store_n(&R[z], (int32_t)(ld_relaxed(&R[z]) | HIGH), __ATOMIC_RELAXED); \
      }                                                                                       \
    }                                                                                         \
    _Pragma("omp barrier")                                                                    \
    _Pragma("omp single")                                                                     \
    t1 = now_ns();                                                                            \
    for (int s = tid; s < T; s += nth) { /* phase 2: moves and in-slice orphans */            \
      const int64_t a = stride * s, b = (a + stride < k) ? a + stride : k;                    \
      int64_t nd = 0, ns = 0;                                                                 \
      for (int64_t i = a; i < b; ++i) {                                                       \
        const uint32_t dd = (uint32_t)R[i];                                                   \
        const int64_t d = (int64_t)(dd & ~HIGH);                                              \
        const int dst_ok = d < sz, src_ok = (dd & HIGH) == 0;                                 \
        if (dst_ok && src_ok) list[d] = list[sz + i];                                         \
        else if (src_ok) list[sz + a + ns++] = list[sz + i];                                  \
        else if (dst_ok) R[a + nd++] = (int32_t)d;                                            \
      }                                                                                       \
      dcnt[s] = nd; scnt[s] = ns;                                                             \
    }                                                                                         \
    _Pragma("omp barrier")                                                                    \
    _Pragma("omp single")                                                                     \
    { /* serial exclusive prefix over slices */                                               \
      int64_t ad = 0, as = 0;                                                                 \
      for (int s = 0; s < T; ++s) {                                                           \
        const int64_t d = dcnt[s], c = scnt[s];                                               \
        dcnt[s] = ad; scnt[s] = as; ad += d; as += c;                                         \
      }                                                                                       \
      dcnt[T] = ad; scnt[T] = as;                                                             \
      t2 = now_ns();                                                                          \
    }                                                                                         \
    /* phase 3: worker s pairs orphan ranks [s*per, (s+1)*per) */                             \
    const int64_t total = dcnt[T] < scnt[T] ? dcnt[T] : scnt[T];                              \
    const int64_t per = (total + T - 1) / T;                                                  \
    for (int s = tid; s < T && per > 0; s += nth) {                                           \
      int64_t t = per * s;                                                                    \
      const int64_t e = (t + per < total) ? t + per : total;                                  \
      if (t >= e) continue;                                                                   \
      int rd = find_row(dcnt, T, t), rs = find_row(scnt, T, t);                               \
      for (; t < e; ++t) {                                                                    \
        while (dcnt[rd + 1] <= t) ++rd;                                                       \
        while (scnt[rs + 1] <= t) ++rs;                                                       \
        const int64_t dst = R[stride * rd + (t - dcnt[rd])];                                  \
        list[dst] = list[sz + stride * rs + (t - scnt[rs])];                                  \
      }                                                                                       \
    }                                                                                         \
  }                                                                                           \
  const int64_t t3 = now_ns();                                                                \
  times[1] = t1 - t0; times[2] = t2 - t1; times[3] = t3 - t2; times[0] = t3 - t0;             \
  free(dcnt); free(scnt);                                                                     \
}

typedef struct { uint64_t a, b; } rzo_u128;
DEFINE_REDZONE_CPU(redzone_cpu_4, uint32_t)
DEFINE_REDZONE_CPU(redzone_cpu_8, uint64_t)
DEFINE_REDZONE_CPU(redzone_cpu_16, rzo_u128)

/* list: n items of esz bytes; R: k distinct indices, clobbered (marks and dest
 * orphans), as in the reference.  threads <= 0 means omp_get_max_threads()-1. */
int rzo_redzone_cpu(void* list, int64_t n, uint32_t esz, int32_t* R, int64_t k, int threads, int64_t times[4]) {
  if (n < 0 || k < 0 || k > n || n >= (1LL << 31)) return RZO_EINVAL;
  if (threads <= 0) { threads = omp_get_max_threads() - 1; if (threads < 1) threads = 1; }
  times[0] = times[1] = times[2] = times[3] = 0;
  if (k == 0 || k == n) return RZO_OK;
  switch (esz) {
    case 4: redzone_cpu_4((uint32_t*)list, n, R, k, threads, times); break;
    case 8: redzone_cpu_8((uint64_t*)list, n, R, k, threads, times); break;
    case 16: redzone_cpu_16((rzo_u128*)list, n, R, k, threads, times); break;
    default: return RZO_EINVAL;
  }
  return RZO_OK;
}

/* ------------------------------------------------------------------------ */
/* O(n) comparator: mark, then stable remove.                               */
/* ------------------------------------------------------------------------ */

/* threads == 1: serial std::remove-style sweep.  threads > 1: parallel mark,
 * parallel per-chunk count, prefix, parallel out-of-place scatter, copy back. */
int rzo_remove_stable(void* data, int64_t n, uint32_t esz, const int32_t* R, int64_t k, int threads, int64_t* time_ns) {
  if (n < 0 || k < 0 || k > n || esz == 0) return RZO_EINVAL;
  char* A = (char*)data;
  uint8_t* mark = (uint8_t*)calloc((size_t)n + 1, 1);
  if (!mark) return RZO_ENOMEM;
  for (int64_t i = 0; i < k; ++i) if (R[i] < 0 || R[i] >= n) { free(mark); return RZO_ERANGE; }
  if (threads < 1) threads = 1;
  const int64_t t0 = now_ns();
  if (threads == 1) {
    for (int64_t i = 0; i < k; ++i) mark[R[i]] = 1;
    int64_t w = 0;
    for (int64_t i = 0; i < n; ++i)
      if (!mark[i]) { if (w != i) memcpy(A + w * esz, A + i * esz, esz); ++w; }
  } else {
    char* tmp = (char*)malloc((size_t)(n - k) * esz + 1);
    int64_t* cnt = (int64_t*)calloc((size_t)threads + 1, sizeof(int64_t));
    if (!tmp || !cnt) { free(tmp); free(cnt); free(mark); return RZO_ENOMEM; }
    const int64_t chunk = (n + threads - 1) / threads;
    _Pragma("omp parallel num_threads(threads)")
    {
      const int tid = omp_get_thread_num();
      _Pragma("omp for schedule(static)")
      for (int64_t i = 0; i < k; ++i) mark[R[i]] = 1;
      const int64_t a = chunk * tid, b = (a + chunk < n) ? a + chunk : n;
      int64_t c = 0;
      for (int64_t i = a; i < b; ++i) c += !mark[i];
      cnt[tid + 1] = c;
      _Pragma("omp barrier")
      _Pragma("omp single")
      for (int s = 0; s < threads; ++s) cnt[s + 1] += cnt[s];
      int64_t w = cnt[tid];
      for (int64_t i = a; i < b; ++i)
        if (!mark[i]) { memcpy(tmp + w * esz, A + i * esz, esz); ++w; }
      _Pragma("omp barrier")
      const int64_t m = n - k, ca = (m + threads - 1) / threads;
      const int64_t lo = ca * tid, hi = (lo + ca < m) ? lo + ca : m;
      if (hi > lo) memcpy(A + lo * esz, tmp + lo * esz, (size_t)(hi - lo) * esz);
    }
    free(tmp); free(cnt);
  }
  if (time_ns) *time_ns = now_ns() - t0;
  free(mark);
  return RZO_OK;
}

/* ------------------------------------------------------------------------ */
/* Invariant checker.                                                       */
/* ------------------------------------------------------------------------ */

/* LSD radix sort of m records of esz bytes (byte-wise, least significant byte
 * first in memory order); result in `a`.  Exact, so multisets compare with memcmp. */
static int radix_sort_bytes(uint8_t* a, int64_t m, uint32_t esz) {
  if (m <= 1) return RZO_OK;
  uint8_t* b = (uint8_t*)malloc((size_t)m * esz);
  if (!b) return RZO_ENOMEM;
  for (uint32_t byte = 0; byte < esz; ++byte) {
    int64_t count[257] = {0};
    for (int64_t i = 0; i < m; ++i) count[a[i * esz + byte] + 1]++;
    for (int d = 0; d < 256; ++d) count[d + 1] += count[d];
    for (int64_t i = 0; i < m; ++i) memcpy(b + (count[a[i * esz + byte]]++) * esz, a + i * esz, esz);
    uint8_t* t = a; a = b; b = t;
  }
  if (esz & 1) { memcpy(b, a, (size_t)m * esz); uint8_t* t = a; a = b; b = t; }
  free(b);
  return RZO_OK;
}

/* Returns a bitmask of violated invariants (0 = all hold):
 *   1  survivor multiset (payload bits) differs from A minus the removed items
 *   2  an item outside the holes moved (in-place invariant: i < n-k, i not in R)
 *   4  exact order differs from a stable remove (mode 1 only)
 *   8  bad input (R out of range or duplicated)
 * mode 0 checks a Redzone result (in-place invariant + multiset of the hole
 * fillers against the surviving tail items, O(n + k log k)); mode 1 checks a
 * stable O(n) result item by item (which implies the multiset).
 * first_bad receives the first offending index when known. */
int rzo_check(const void* orig, const void* out, int64_t n, uint32_t esz, const int32_t* R, int64_t k,
              int mode, int64_t* first_bad) {
  const uint8_t* O = (const uint8_t*)orig;
  const uint8_t* X = (const uint8_t*)out;
  int bad = 0;
  int64_t fb = -1;
  uint8_t* inR = (uint8_t*)calloc((size_t)n + 1, 1);
  if (!inR) return 16;
  for (int64_t i = 0; i < k; ++i) {
    if (R[i] < 0 || R[i] >= n || inR[R[i]]) { free(inR); if (first_bad) *first_bad = i; return 8; }
    inR[R[i]] = 1;
  }
  const int64_t sz = n - k;
  if (mode == 1) {
    int64_t w = 0;
    for (int64_t i = 0; i < n; ++i) {
      if (inR[i]) continue;
      if (memcmp(X + w * esz, O + i * esz, esz) != 0) { bad |= 4; fb = w; break; }
      ++w;
    }
    free(inR);
    if (first_bad) *first_bad = fb;
    return bad;
  }
  int64_t h = 0;
  for (int64_t i = 0; i < sz; ++i) {
    if (inR[i]) { ++h; continue; }
    if (memcmp(X + i * esz, O + i * esz, esz) != 0) { bad |= 2; if (fb < 0) fb = i; }
  }
  uint8_t* got = (uint8_t*)malloc((size_t)(h ? h : 1) * esz);
  uint8_t* want = (uint8_t*)malloc((size_t)(h ? h : 1) * esz);
  if (!got || !want) { free(got); free(want); free(inR); return 16; }
  int64_t g = 0, w = 0;
  for (int64_t i = 0; i < sz; ++i) if (inR[i]) memcpy(got + (g++) * esz, X + i * esz, esz);
  for (int64_t i = sz; i < n; ++i) if (!inR[i]) { if (w < h) memcpy(want + w * esz, O + i * esz, esz); ++w; }
  if (w != h) bad |= 1;
  else {
    radix_sort_bytes(got, h, esz);
    radix_sort_bytes(want, h, esz);
    if (memcmp(got, want, (size_t)h * esz) != 0) bad |= 1;
  }
  free(got); free(want); free(inR);
  if (first_bad) *first_bad = fb;
  return bad;
}

/* Plain multiset test over the whole survivor range [0, n-k), for any
 * arrangement: 0 when it equals A minus the removed items, payload bits included. */
int rzo_check_multiset(const void* orig, const void* out, int64_t n, uint32_t esz, const int32_t* R, int64_t k) {
  const uint8_t* O = (const uint8_t*)orig;
  const uint8_t* X = (const uint8_t*)out;
  uint8_t* inR = (uint8_t*)calloc((size_t)n + 1, 1);
  if (!inR) return 16;
  for (int64_t i = 0; i < k; ++i) {
    if (R[i] < 0 || R[i] >= n || inR[R[i]]) { free(inR); return 8; }
    inR[R[i]] = 1;
  }
  const int64_t m = n - k;
  uint8_t* got = (uint8_t*)malloc((size_t)(m ? m : 1) * esz);
  uint8_t* want = (uint8_t*)malloc((size_t)(m ? m : 1) * esz);
  if (!got || !want) { free(got); free(want); free(inR); return 16; }
  memcpy(got, X, (size_t)m * esz);
  int64_t w = 0;
  for (int64_t i = 0; i < n; ++i) if (!inR[i]) memcpy(want + (w++) * esz, O + i * esz, esz);
  radix_sort_bytes(got, m, esz);
  radix_sort_bytes(want, m, esz);
  const int bad = memcmp(got, want, (size_t)m * esz) != 0;
  free(got); free(want); free(inR);
  return bad;
}
